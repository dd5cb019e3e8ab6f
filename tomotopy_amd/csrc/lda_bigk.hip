// lda_bigk.hip -- gfx950 sampler kernels for K above 2048 topics (up to
// tomotopy's 32767, LDAModel.hpp:964-984 with Tid = uint16).
//
// The K <= 2048 kernels keep each lane's KPL <= 32 topics of every K-vector in
// registers or LDS.  Beyond that the per-document state no longer fits, so
// these kernels keep it in device memory and walk each lane's topics in
// chunks of 16:
//   * n_dk is the document's row of ndk itself (one wavefront owns a document
//     while it samples it, in every schedule); the private n_k of a worker is
//     its row of per-worker counts; the current word's column is the n_kw row
//     itself where the schedule gives this worker the cells exclusively
//     (none / partition / copy_merge, fold-in's private run columns) and a
//     per-wavefront copy where n_kw is shared (the device-wide sweep, whose
//     cell updates go to n_kw as atomics).
//   * Every position of a K-vector is read and written by one lane only (the
//     lane-layout owner of that topic), so a lane's own program order orders
//     its count updates before its next reads; no cross-lane traffic goes
//     through memory.
//   * The lane layout is the K <= 2048 one (topic k = l*KPL + j at
//     (j/4)*256 + 4l + j%4), so chunk c of every lane is the contiguous block
//     of positions [1024c, 1024(c+1)) and the host's layout conversions,
//     read-outs, LL and alpha kernels apply unchanged.
// The arithmetic is the K <= 2048 sampler's: weights
// ((n_dk+alpha)(n_kw+eta))/(n_k+V*eta) with the Markstein quotient, each
// lane's running sum over its topics in order (chunk after chunk), the DPP
// wave scan of the lane totals, and the count of a lane's cdf values <= u*total.
// The per-chunk last values of the running sum sit in LDS; the search finds
// the first chunk whose last cdf value exceeds r and recomputes that chunk's
// 16 weights (the same operations on the same counts, so the same values) to
// count inside it.  The result is the K <= 2048 sampler's draw for the same
// state: the oracle's device-order scan covers every topics-per-lane count.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "lda_internal.h"
#include "lda_kernels_common.h"

namespace lda {
namespace {

// CH topics per lane per chunk (16, or KPL itself when it is smaller); chunk
// c of every lane is the position block [64*CH*c, 64*CH*(c+1))
constexpr int BK_MAX_CHUNKS = kBigMaxKpl / 16;

template <bool WT>
using bk_t = typename std::conditional<WT, float, int32_t>::type;

__device__ __forceinline__ void bk_lane_limits(int K, int lane, int kpl, int& nvalid, int& lastLane) {
    const int nv = K - lane * kpl;
    nvalid = nv < 0 ? 0 : (nv > kpl ? kpl : nv);
    lastLane = (K - 1) / kpl;
}

// eta of the current word and the per-topic eta sums, one chunk at a time
// (EtaHelper, LDAModel.hpp:81-117).  Sym: eta and V*eta.  Prior: the word's
// eta column (a row of the sparse prior table, or eta) and sum_v eta_kv.
struct BkEtaSym {
    float eta, veta;
    __device__ __forceinline__ void word(int) {}
    template <int CH>
    __device__ __forceinline__ void chunk(int, int, fvec<CH>& e, fvec<CH>& s) const {
        e = eta;
        s = veta;
    }
};

struct BkEtaPrior {
    const int32_t* prior_row;
    const float* prior_eta;
    const float* etasum;
    const float* ecol;  // current word's eta column (NULL: eta everywhere)
    float eta;
    int Kp;
    __device__ __forceinline__ void word(int v) {
        const int r = prior_row[v];
        ecol = r >= 0 ? prior_eta + (int64_t)r * Kp : nullptr;
    }
    template <int CH>
    __device__ __forceinline__ void chunk(int c, int lane, fvec<CH>& e, fvec<CH>& s) const {
        if (ecol) load_kvec_f<CH>(ecol + c * 64 * CH, lane, e);
        else e = eta;
        load_kvec_f<CH>(etasum + c * 64 * CH, lane, s);
    }
};

template <bool ASYM>
struct BkEtaOf { using type = BkEtaSym; };
template <>
struct BkEtaOf<true> { using type = BkEtaPrior; };

template <bool ASYM>
__device__ __forceinline__ typename BkEtaOf<ASYM>::type bk_make_eta(const int32_t* prior_row, const float* prior_eta,
                                                                  const float* etasum, float eta, float veta, int Kp) {
    if constexpr (ASYM) return BkEtaPrior{prior_row, prior_eta, etasum, nullptr, eta, Kp};
    else return BkEtaSym{eta, veta};
}

// One chunk of this lane's state (CH topics of n_dk, the column, n_k, alpha
// and the eta terms) in registers, loaded apart from its use so that the
// next chunk's loads are in flight while this one is weighed.
template <int CH, bool WT>
struct BkChunk {
    using V = typename std::conditional<WT, fvec<CH>, ivec<CH>>::type;
    V nd, col, nk;
    fvec<CH> al, e, s;

    template <class E>
    __device__ __forceinline__ void load(const bk_t<WT>* ndp, const bk_t<WT>* colp, const bk_t<WT>* nkp,
                                         const float* alpha, int c, int lane, const E& et) {
        const int o = c * (64 * CH);
        load_kvec_f<CH>(alpha + o, lane, al);
        et.template chunk<CH>(c, lane, e, s);
        if constexpr (WT) {
            load_kvec_f<CH>(ndp + o, lane, nd);
            load_kvec_f<CH>(colp + o, lane, col);
            load_kvec_f<CH>(nkp + o, lane, nk);
        } else {
            load_kvec<CH>(ndp + o, lane, nd);
            load_kvec<CH>(colp + o, lane, col);
            load_kvec<CH>(nkp + o, lane, nk);
        }
    }

    // The weights (getZLikelihoods, LDAModel.hpp:273-285) in the K <= 2048
    // sampler's operations: integer counts form A = float(n_dk) + alpha,
    // B = float(n_kw) + eta, b = float(n_k) + V*eta; float counts
    // (n_dk + alpha)(n_kw + eta) and n_k + V*eta; then q = a/b as a*y with
    // one Markstein correction, y = RN(1/b).
    __device__ __forceinline__ void weights(fvec<CH>& w) const {
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            float a, bf;
            if constexpr (WT) {
                a = (nd[j] + al[j]) * (col[j] + e[j]);
                bf = nk[j] + s[j];
            } else {
                const float A = (float)nd[j] + al[j];
                const float B = (float)col[j] + e[j];
                bf = (float)nk[j] + s[j];
                a = A * B;
            }
            const float y = recip_rn(bf);
            const float q0 = a * y;
            const float rr = __builtin_fmaf(-q0, bf, a);
            w[j] = __builtin_fmaf(rr, y, q0);
        }
    }
};

// addWordTo for topic z (LDAModel.hpp:287-299) by the owner lane: +-1 for
// integer counts; float counts add w or take max(x - w, 0) (updateCnt,
// Utils.hpp:132-155).
// SHARED: n_k is the workgroup's, in LDS, shared by its wavefronts (LDS
// atomics; a wavefront reading it may see another's update of a moment ago,
// the K <= 2048 device-wide sweep's rule).
template <bool WT, bool SHARED = false>
__device__ __forceinline__ void bk_bump(bk_t<WT>* nd, bk_t<WT>* col, bk_t<WT>* nk, int lane, int kpl, int ksh, int z,
                                        float w, bool dec) {
    if (lane == (z >> ksh)) {
        // the three cells are read before any is written: the compiler cannot
        // tell the arrays apart, and interleaved read-modify-writes would
        // cost three memory round trips instead of one
        const int q = kpos_rt(kpl, z);
        const bk_t<WT> a = nd[q], c = col[q];
        if constexpr (WT) {
            if constexpr (SHARED) {
                nd[q] = dec ? sub_clamp(a, w) : a + w;
                col[q] = dec ? sub_clamp(c, w) : c + w;
                if (dec) atomic_sub_clamp(nk + q, w);
                else atomicAdd(nk + q, w);
            } else {
                const float n = nk[q];
                nd[q] = dec ? sub_clamp(a, w) : a + w;
                col[q] = dec ? sub_clamp(c, w) : c + w;
                nk[q] = dec ? sub_clamp(n, w) : n + w;
            }
        } else {
            const int d = dec ? -1 : 1;
            if constexpr (SHARED) {
                nd[q] = a + d;
                col[q] = c + d;
                atomicAdd(nk + q, d);
            } else {
                const int n = nk[q];
                nd[q] = a + d;
                col[q] = c + d;
                nk[q] = n + d;
            }
        }
    }
}

// Running sum of the chunk's weights from `carry` (the lane's earlier chunks).
template <int CH>
__device__ __forceinline__ void bk_prefix(fvec<CH>& w, float carry) {
    w[0] = carry + w[0];
#pragma unroll
    for (int j = 1; j < CH; ++j) w[j] = w[j - 1] + w[j];
}

// One token: remove zo, weights, scan, search, add the new topic (returned,
// wave-uniform).  lastc: this wavefront's LDS row of chunk last values
// ([chunk][lane]).  cdf: NULL, or this wavefront's LDS copy of every running
// sum (Kp floats, lane layout), from which the search reads the chunk it
// needs instead of weighing it again.  *total receives cdf[K-1].
template <int CH, bool WT, class E, bool SHARED = false>
__device__ __forceinline__ int bk_token(bk_t<WT>* nd, bk_t<WT>* col, bk_t<WT>* nk, const float* alpha, int lane,
                                        int kpl, int ksh, int zo, float tw, float u, int K, int nvalid, int lastLane,
                                        const E& et, float* lastc, float* cdf, float* total) {
    bk_bump<WT, SHARED>(nd, col, nk, lane, kpl, ksh, zo, tw, true);
    const int C = kpl / CH;
    float carry = 0.0f;
    // chunks in pairs: the next chunk's loads are issued before this one is weighed
    BkChunk<CH, WT> A, B;
    A.load(nd, col, nk, alpha, 0, lane, et);
    for (int c = 0; c < C; c += 2) {
        if (c + 1 < C) B.load(nd, col, nk, alpha, c + 1, lane, et);
        {
            fvec<CH> w;
            A.weights(w);
            bk_prefix<CH>(w, carry);
            carry = w[CH - 1];
            lastc[c * 64 + lane] = carry;
            if (cdf) store_kvec<CH>(reinterpret_cast<int*>(cdf + c * (64 * CH)), lane, __builtin_bit_cast(ivec<CH>, w));
        }
        if (c + 2 < C) A.load(nd, col, nk, alpha, c + 2, lane, et);
        if (c + 1 < C) {
            fvec<CH> w;
            B.weights(w);
            bk_prefix<CH>(w, carry);
            carry = w[CH - 1];
            lastc[(c + 1) * 64 + lane] = carry;
            if (cdf)
                store_kvec<CH>(reinterpret_cast<int*>(cdf + (c + 1) * (64 * CH)), lane, __builtin_bit_cast(ivec<CH>, w));
        }
    }
    const float incl = wave_inclusive_scan(carry);
    const float excl = dpp_wave_shr1(incl);
    const float last = excl + carry;
    const float tot = readlane_f(last, lastLane);
    const float r = u * tot;
    // the lane's cdf values are non-decreasing, so the chunks whose last cdf
    // value is <= r form a prefix; count inside the first chunk beyond it
    int cs = 0;
    for (int c = 0; c < C; ++c) cs += (r < excl + lastc[c * 64 + lane]) ? 0 : 1;
    int miss = kpl;
    if (cs < C) {
        fvec<CH> w;
        if (cdf) {
            load_kvec_f<CH>(cdf + cs * (64 * CH), lane, w);
        } else {
            A.load(nd, col, nk, alpha, cs, lane, et);
            A.weights(w);
            bk_prefix<CH>(w, cs ? lastc[(cs - 1) * 64 + lane] : 0.0f);
        }
        miss = cs * CH + cdf_count_le<CH>(w, excl, r, excl + w[CH - 1]);
    }
    const unsigned long long m = __ballot(miss < nvalid);
    int zn;
    if (m == 0ull) zn = K - 1;
    else {
        const int L = __builtin_ctzll(m);
        zn = L * kpl + readlane_i(miss, L);
    }
    zn = uniform_i(zn);
    bk_bump<WT, SHARED>(nd, col, nk, lane, kpl, ksh, zn, tw, false);
    if (total) *total = tot;
    return zn;
}

// This lane's positions of a K-vector: dst = src.
template <int CH, typename T>
__device__ __forceinline__ void bk_copy(T* dst, const T* src, int lane, int kpl) {
    for (int c = 0; c < kpl / CH; ++c) {
        ivec<CH> x;
        load_kvec<CH>(reinterpret_cast<const int32_t*>(src) + c * (64 * CH), lane, x);
        store_kvec<CH>(reinterpret_cast<int32_t*>(dst) + c * (64 * CH), lane, x);
    }
}

template <int CH, typename T>
__device__ __forceinline__ void bk_fill(T* dst, int lane, int kpl) {
    ivec<CH> z = 0;
    for (int c = 0; c < kpl / CH; ++c) store_kvec<CH>(reinterpret_cast<int32_t*>(dst) + c * (64 * CH), lane, z);
}

// sampleDocument over the token range [b, e) (LDAModel.hpp:326-354) on the
// state (nd, nk).  MODE_DET: the column is the n_kw row itself (this worker
// owns its cells); MODE_ATOMIC: colbuf holds this wavefront's copy of the
// column (read when the word's run starts, plus its own updates) and the
// token's net change goes to the shared n_kw as atomics.
template <int CH, int MODE, bool WT, class E, bool SHARED = false>
__device__ __forceinline__ void bk_sample_range(bk_t<WT>* nd, bk_t<WT>* nk, bk_t<WT>* colbuf, const float* alpha,
                                                const SweepParams& p, uint32_t sweep, int lane, uint64_t b, uint64_t e,
                                                int kpl, int ksh, int nvalid, int lastLane, E& et, float* lastc,
                                                float* cdf = nullptr) {
    using T = bk_t<WT>;
    T* nkwT = reinterpret_cast<T*>(p.nkw);
    // MODE_ATOMIC: the n_kw changes of each group of col_group tokens go out at
    // the group's end as whole-wavefront atomics, lane t those of token t (the
    // K <= 2048 sweep's rule: per-token single-lane atomics would hold the
    // next column copy's loads behind their round trips)
    int TF = 64;
    if constexpr (MODE == MODE_ATOMIC) {
        const int cg = p.col_group;
        TF = (cg >= 1 && cg <= 64 && (cg & (cg - 1)) == 0) ? cg : 64;
    }
    int vcur = -1;
    T* col = nullptr;
    // the current block of up to 64 tokens, token t of it on lane t
    int nb = 0, wv = -1, zv = 0xFFFF, zkeep = 0xFFFF;
    float uv = 0.0f, tw = 1.0f;
    uint64_t g = 0;
    for (uint64_t t0 = b; t0 < e; t0 += TF) {
        const int off = (int)((t0 - b) & 63);
        if (off == 0) {
            nb = (int)((e - t0) < 64 ? (e - t0) : 64);
            g = t0 + (uint64_t)lane;
            wv = -1;
            zv = 0xFFFF;
            uv = 0.0f;
            tw = 1.0f;
            if (lane < nb) {
                wv = (int)p.words[g];
                zv = p.z[g];
                if constexpr (WT) tw = p.wt[g];
                uv = uniform24(philox_x0(p.seed_lo, p.seed_hi, (uint32_t)g, (uint32_t)(g >> 32), sweep, TAG_SWEEP));
            }
            zkeep = zv;
        }
        const int i1 = off + TF < nb ? off + TF : nb;
        for (int i = off; i < i1; ++i) {
            const int v = readlane_i(wv, i);
            if ((unsigned)v >= (unsigned)p.V) continue;
            if (v != vcur) {
                if constexpr (MODE == MODE_DET) {
                    col = nkwT + (int64_t)v * p.Kp;
                } else {
                    bk_copy<CH>(colbuf, nkwT + (int64_t)v * p.Kp, lane, kpl);
                    col = colbuf;
                }
                et.word(v);
                vcur = v;
            }
            const int zo = readlane_i(zv, i);
            const float w = readlane_f(tw, i);
            const float u = readlane_f(uv, i);
            const int zn = bk_token<CH, WT, E, SHARED>(nd, col, nk, alpha, lane, kpl, ksh, zo, w, u, p.K, nvalid, lastLane,
                                                       et, lastc, cdf, nullptr);
            zkeep = (lane == i) ? zn : zkeep;
        }
        if constexpr (MODE == MODE_ATOMIC) {
            if (lane >= off && lane < i1 && zkeep != zv && (unsigned)wv < (unsigned)p.V) {
                T* cp = nkwT + (int64_t)wv * p.Kp;
                const int qo = kpos_rt(kpl, zv), qn = kpos_rt(kpl, zkeep);
                if constexpr (WT) {
                    // the merge's clamp max(n_kw, 0) removes rounding-level negatives
                    unsafeAtomicAdd(cp + qo, -tw);
                    unsafeAtomicAdd(cp + qn, tw);
                } else {
                    atomicAdd(cp + qo, -1);
                    atomicAdd(cp + qn, 1);
                }
            }
        }
        if (i1 == nb && lane < nb) p.z[g] = (uint16_t)zkeep;
    }
}

// The one-wavefront deterministic sweeps keep the document's n_dk row and the
// worker's n_k in LDS as well up to 128 topics per lane (98 KB with the
// running sums at 128).
__host__ __device__ constexpr bool bk_det_rows(int kpl) { return kpl <= 128; }

// Dynamic LDS of the kernels below.  One-wavefront kernels: the running sums
// (Kp floats, when kept) then the chunk last values (64 per chunk).  The
// sweeps keep the sums at every K (a few wavefronts run at once); the kernels
// that fill the chip with one document per wavefront keep them up to 64 topics
// per lane, where they still leave room for several wavefronts per CU.
extern __shared__ int bk_lds[];
constexpr int kBkCdfManyMaxKpl = 64;

__device__ __forceinline__ uint64_t bk_prime(uint64_t N, uint32_t sseed) {
    uint64_t P = 1;
    if (N) {
        P = kShufflePrimes[sseed & 0xF];
        if (N % P == 0) {
            P = kShufflePrimes[(sseed + 1) & 0xF];
            if (N % P == 0) P = kShufflePrimes[(sseed + 2) & 0xF];
            if (N % P == 0) P = kShufflePrimes[(sseed + 3) & 0xF];
        }
        P %= N;
    }
    return P;
}

// One round of the partition schedule (LDAModel.hpp:373-399), `none` for
// W == 1: worker w = blockIdx.x visits documents id*W + (round+w) % W in
// forShuffled order (Utils.hpp:157-178), their tokens in vocabulary chunk w;
// private n_k in nk_out row w.
template <int CH, bool WT, bool ASYM>
__global__ void __launch_bounds__(64) bk_sweep_det_kernel(SweepParams p, int kpl, int ksh) {
    using T = bk_t<WT>;
    float* cdf = reinterpret_cast<float*>(bk_lds);
    float* lastc = cdf + p.Kp;
    // with rows (bk_det_rows): the document's n_dk row and the worker's n_k
    // in LDS after the chunk last values, staged in and written back
    T* ndl = reinterpret_cast<T*>(lastc + (kpl / CH) * 64);
    T* nkl = ndl + p.Kp;
    const bool rows = bk_det_rows(kpl);
    const int lane = lane_id();
    const int w = blockIdx.x;
    int nvalid, lastLane;
    bk_lane_limits(p.K, lane, kpl, nvalid, lastLane);
    auto et = bk_make_eta<ASYM>(p.prior_row, p.prior_eta, p.etasum, p.eta, p.veta, p.Kp);
    T* nkg = reinterpret_cast<T*>(p.nk_out) + (int64_t)w * p.Kp;
    if (rows) bk_copy<CH>(nkl, nkg, lane, kpl);
    T* nkp = rows ? nkl : nkg;
    const int W = p.W;
    const uint32_t sweep = p.sweep_dev ? (uint32_t)uniform_i((int)*p.sweep_dev) : p.sweep;
    const int didx = (p.round + w) % W;
    const uint64_t N = (uint64_t)((p.D + (W - 1) - didx) / W);
    const uint32_t sseed = philox_x0(p.seed_lo, p.seed_hi, (uint32_t)w, (uint32_t)p.round, sweep, TAG_SHUFFLE);
    const uint64_t P = bk_prime(N, sseed);
    for (uint64_t i = 0; i < N; ++i) {
        const uint64_t id = ((i + (uint64_t)sseed) * P) % N;
        const int64_t d = (int64_t)id * W + didx;
        const uint64_t b = p.coff[(int64_t)w * p.D + d], e = p.coff[(int64_t)(w + 1) * p.D + d];
        if (b == e) continue;
        T* nd = reinterpret_cast<T*>(p.ndk) + d * p.Kp;
        if (rows) bk_copy<CH>(ndl, nd, lane, kpl);
        bk_sample_range<CH, MODE_DET, WT>(rows ? ndl : nd, nkp, nullptr, p.alpha, p, sweep, lane, b, e, kpl, ksh, nvalid,
                                          lastLane, et, lastc, cdf);
        if (rows) bk_copy<CH>(nd, ndl, lane, kpl);
    }
    if (rows) bk_copy<CH>(nkg, nkl, lane, kpl);
}

// copy_merge (LDAModel.hpp:401-423): worker w samples on its private n_kw
// (nkw + w*V*Kp) and n_k (nk_out row w) the tasks c = w, w+W, ... of
// min(8W, D) strided tasks, each in forShuffled order.
template <int CH, bool WT, bool ASYM>
__global__ void __launch_bounds__(64) bk_sweep_copy_kernel(SweepParams p, int kpl, int ksh) {
    using T = bk_t<WT>;
    float* cdf = reinterpret_cast<float*>(bk_lds);
    float* lastc = cdf + p.Kp;
    T* ndl = reinterpret_cast<T*>(lastc + (kpl / CH) * 64);  // rows in LDS, as in bk_sweep_det_kernel
    T* nkl = ndl + p.Kp;
    const bool rows = bk_det_rows(kpl);
    const int lane = lane_id();
    const int w = blockIdx.x;
    int nvalid, lastLane;
    bk_lane_limits(p.K, lane, kpl, nvalid, lastLane);
    auto et = bk_make_eta<ASYM>(p.prior_row, p.prior_eta, p.etasum, p.eta, p.veta, p.Kp);
    T* nkg = reinterpret_cast<T*>(p.nk_out) + (int64_t)w * p.Kp;
    if (rows) bk_copy<CH>(nkl, nkg, lane, kpl);
    T* nkp = rows ? nkl : nkg;
    SweepParams q = p;
    q.nkw = p.nkw + (int64_t)w * p.V * p.Kp;
    const int W = p.W;
    const int64_t chStride = (int64_t)W * 8 < p.D ? (int64_t)W * 8 : p.D;
    for (int64_t ch = w; ch < chStride; ch += W) {
        const uint64_t N = (uint64_t)((p.D + (chStride - 1) - ch) / chStride);
        const uint32_t sseed = philox_x0(p.seed_lo, p.seed_hi, (uint32_t)ch, 0xC0u, p.sweep, TAG_SHUFFLE);
        const uint64_t P = bk_prime(N, sseed);
        for (uint64_t i = 0; i < N; ++i) {
            const int64_t d = (int64_t)(((i + (uint64_t)sseed) * P) % N) * chStride + ch;
            const uint64_t b = p.off[d], e = p.off[d + 1];
            if (b == e) continue;
            T* nd = reinterpret_cast<T*>(p.ndk) + d * p.Kp;
            if (rows) bk_copy<CH>(ndl, nd, lane, kpl);
            bk_sample_range<CH, MODE_DET, WT>(rows ? ndl : nd, nkp, nullptr, p.alpha, q, q.sweep, lane, b, e, kpl, ksh,
                                              nvalid, lastLane, et, lastc, cdf);
            if (rows) bk_copy<CH>(nd, ndl, lane, kpl);
        }
    }
    if (rows) bk_copy<CH>(nkg, nkl, lane, kpl);
}

// Device-wide sweep: every wavefront is a worker with a private n_k (the
// sweep-start snapshot plus its own changes) and a private copy of the
// current column, in bk_scratch rows 2*blockIdx.x and 2*blockIdx.x + 1;
// documents come longest first from a global counter; n_kw takes the
// changes as atomics.  Integer counts add each worker's n_k change to nk_out
// at the end; float counts get n_k from the merge's row sums.
template <int CH, bool WT, bool ASYM>
__global__ void __launch_bounds__(64) bk_sweep_device_kernel(SweepParams p, int kpl, int ksh) {
    using T = bk_t<WT>;
    __shared__ float lastc[BK_MAX_CHUNKS * 64];
    const int lane = lane_id();
    int nvalid, lastLane;
    bk_lane_limits(p.K, lane, kpl, nvalid, lastLane);
    auto et = bk_make_eta<ASYM>(p.prior_row, p.prior_eta, p.etasum, p.eta, p.veta, p.Kp);
    T* nkp = reinterpret_cast<T*>(p.bk_scratch) + (int64_t)blockIdx.x * 2 * p.Kp;
    T* colbuf = nkp + p.Kp;
    const T* snap = reinterpret_cast<const T*>(p.nk_in);
    bk_copy<CH>(nkp, snap, lane, kpl);
    for (;;) {
        uint32_t idx = 0;
        if (lane == 0) idx = atomicAdd(p.counter, 1u);
        idx = (uint32_t)uniform_i((int)idx);
        if ((int64_t)idx >= p.D) break;
        const uint32_t d = p.doc_order[idx];
        T* nd = reinterpret_cast<T*>(p.ndk) + (int64_t)d * p.Kp;
        bk_sample_range<CH, MODE_ATOMIC, WT>(nd, nkp, colbuf, p.alpha, p, p.sweep, lane, p.off[d], p.off[d + 1], kpl,
                                             ksh, nvalid, lastLane, et, lastc);
    }
    if constexpr (!WT) {
        for (int c = 0; c < kpl / CH; ++c) {
            ivec<CH> a, s;
            load_kvec<CH>(nkp + c * (64 * CH), lane, a);
            load_kvec<CH>(snap + c * (64 * CH), lane, s);
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                const int dlt = a[j] - s[j];
                if (dlt) atomicAdd(p.nk_out + kpos_rt(kpl, lane * kpl + c * CH + j), dlt);
            }
        }
    }
}

// The same schedule with workgroups of several wavefronts sharing, in LDS,
// alpha and one private n_k (the snapshot plus the workgroup's changes, LDS
// atomics), so that per token only the n_dk row and the wavefront's column
// copy (bk_scratch row blockIdx.x * waves + wave) come from device memory.
// ROWS: those two rows are in LDS too (the n_dk row staged in at the
// document's start and written back at its end; each lane copies and uses
// only its own positions, so no barrier is needed).
// Dynamic LDS: alpha (Kp) | n_k (Kp) | per wavefront the chunk last values
// (and with ROWS its n_dk row and column copy, Kp each).

template <int CH, bool WT, bool ASYM, int ROWS>
__global__ void __launch_bounds__(64 * kMaxSweepWaves) bk_sweep_device_lds_kernel(SweepParams p, int kpl, int ksh) {
    using T = bk_t<WT>;
    const int lane = lane_id();
    const int wave = (int)(threadIdx.x >> 6), waves = (int)(blockDim.x >> 6);
    const int Kp = p.Kp;
    const int lc = (kpl / CH) * 64;
    float* al = reinterpret_cast<float*>(bk_lds);
    T* nkl = reinterpret_cast<T*>(bk_lds + Kp);
    constexpr bool ND_L = (ROWS & 1) != 0, COL_L = (ROWS & 2) != 0;
    float* lastc = reinterpret_cast<float*>(bk_lds + 2 * Kp) + wave * (lc + (ND_L ? Kp : 0) + (COL_L ? Kp : 0));
    T* ndl = reinterpret_cast<T*>(lastc + lc);
    T* coll = ndl + (ND_L ? Kp : 0);
    int nvalid, lastLane;
    bk_lane_limits(p.K, lane, kpl, nvalid, lastLane);
    auto et = bk_make_eta<ASYM>(p.prior_row, p.prior_eta, p.etasum, p.eta, p.veta, Kp);
    const T* snap = reinterpret_cast<const T*>(p.nk_in);
    for (int q = (int)threadIdx.x; q < Kp; q += (int)blockDim.x) {
        al[q] = p.alpha[q];
        nkl[q] = snap[q];
    }
    __syncthreads();
    T* colbuf = COL_L ? coll : reinterpret_cast<T*>(p.bk_scratch) + ((int64_t)blockIdx.x * waves + wave) * Kp;
    for (;;) {
        uint32_t idx = 0;
        if (lane == 0) idx = atomicAdd(p.counter, 1u);
        idx = (uint32_t)uniform_i((int)idx);
        if ((int64_t)idx >= p.D) break;
        const uint32_t d = p.doc_order[idx];
        T* nd = reinterpret_cast<T*>(p.ndk) + (int64_t)d * Kp;
        if constexpr (ND_L) {
            bk_copy<CH>(ndl, nd, lane, kpl);
            bk_sample_range<CH, MODE_ATOMIC, WT, decltype(et), true>(ndl, nkl, colbuf, al, p, p.sweep, lane, p.off[d],
                                                                     p.off[d + 1], kpl, ksh, nvalid, lastLane, et,
                                                                     lastc);
            bk_copy<CH>(nd, ndl, lane, kpl);
        } else {
            bk_sample_range<CH, MODE_ATOMIC, WT, decltype(et), true>(nd, nkl, colbuf, al, p, p.sweep, lane, p.off[d],
                                                                     p.off[d + 1], kpl, ksh, nvalid, lastLane, et,
                                                                     lastc);
        }
    }
    if constexpr (!WT) {
        __syncthreads();
        for (int q = (int)threadIdx.x; q < Kp; q += (int)blockDim.x) {
            const int dlt = nkl[q] - snap[q];
            if (dlt) atomicAdd(p.nk_out + q, dlt);
        }
    }
}

// Frozen single steps: for token t the state (n_dk of doc[t], the column of
// word[t], n_k) is copied into this wavefront's scratch rows, the token is
// removed, sampled with u[t] and added back.
template <int CH>
__global__ void __launch_bounds__(64) bk_sample_frozen_kernel(const int32_t* __restrict__ ndk,
                                                              const int32_t* __restrict__ nkw,
                                                              const int32_t* __restrict__ nk, const float* alpha,
                                                              float eta, float veta, int K, int Kp, int64_t n,
                                                              const int64_t* doc, const uint32_t* word,
                                                              const uint16_t* z_old, const float* u, uint16_t* z_new,
                                                              float* total, int kpl, int ksh, int32_t* scratch) {
    float* cdf = kpl <= kBkCdfManyMaxKpl ? reinterpret_cast<float*>(bk_lds) : nullptr;
    float* lastc = reinterpret_cast<float*>(bk_lds) + (cdf ? Kp : 0);
    const int lane = lane_id();
    int nvalid, lastLane;
    bk_lane_limits(K, lane, kpl, nvalid, lastLane);
    const BkEtaSym et{eta, veta};
    int32_t* nd = scratch + (int64_t)blockIdx.x * 3 * Kp;
    int32_t* col = nd + Kp;
    int32_t* nkc = col + Kp;
    for (int64_t t = blockIdx.x; t < n; t += gridDim.x) {
        bk_copy<CH>(nd, ndk + doc[t] * Kp, lane, kpl);
        bk_copy<CH>(col, nkw + (int64_t)word[t] * Kp, lane, kpl);
        bk_copy<CH>(nkc, nk, lane, kpl);
        float tot;
        const int zn = bk_token<CH, false>(nd, col, nkc, alpha, lane, kpl, ksh, (int)z_old[t], 1.0f, u[t], K, nvalid,
                                       lastLane, et, lastc, cdf, &tot);
        if (lane == 0) {
            z_new[t] = (uint16_t)zn;
            if (total) total[t] = tot;
        }
    }
}

// Fold-in (together=False, LDAModel.hpp:920-953) of one document per
// wavefront, documents from a counter (longest first) or job job_base alone.
// The document's private state: n_dk in ndk_out row slot, n_k of the private
// copy in nk_out row slot, and for each run of equal word ids its private
// column in `cells` (initialised from the model's column).  Initial topics
// from Discrete(alpha) (or the word's eta column with priors), added in token
// order; then `iters` sequential sweeps.  Integer counts reproduce the
// integer fold-in (n_kw + c_d per run); float counts follow tomoto's clamped
// updates in token order.
// ROWS: the document's n_dk and n_k rows live in LDS while it is sampled
// (after the running sums and chunk last values) and are written to
// ndk_out / nk_out at its end; the run columns stay in device memory.
template <int CH, bool WT, bool ASYM, bool ROWS>
__global__ void __launch_bounds__(64) bk_infer_gen_kernel(InferGenParams p, int kpl, int ksh) {
    using T = bk_t<WT>;
    float* cdf = kpl <= kBkCdfManyMaxKpl ? reinterpret_cast<float*>(bk_lds) : nullptr;
    float* lastc = reinterpret_cast<float*>(bk_lds) + (cdf ? p.Kp : 0);
    T* ndl = reinterpret_cast<T*>(lastc + (kpl / CH) * 64);
    T* nkl = ndl + p.Kp;
    const int lane = lane_id();
    int nvalid, lastLane;
    bk_lane_limits(p.K, lane, kpl, nvalid, lastLane);
    auto et = bk_make_eta<ASYM>(p.prior_row, p.prior_eta, p.etasum, p.eta, p.veta, p.Kp);
    const T* nkwT = reinterpret_cast<const T*>(p.nkw);
    T* cells = reinterpret_cast<T*>(p.cells);
    for (;;) {
        int64_t idx;
        if (!p.counter) {
            idx = p.job_base;
        } else {
            uint32_t t = 0;
            if (lane == 0) t = atomicAdd(p.counter, 1u);
            idx = (int64_t)(uint32_t)uniform_i((int)t);
        }
        if (idx >= p.njobs) break;
        const InferJob J = p.jobs[idx];
        const uint64_t b = J.b, e = J.e;
        T* ndg = reinterpret_cast<T*>(p.ndk_out) + (int64_t)J.slot * p.Kp;
        T* nkg = reinterpret_cast<T*>(p.nk_out) + (int64_t)J.slot * p.Kp;
        T* nd = ROWS ? ndl : ndg;
        T* nk = ROWS ? nkl : nkg;
        bk_fill<CH>(nd, lane, kpl);
        bk_copy<CH>(nk, reinterpret_cast<const T*>(p.nk), lane, kpl);
        // initial topics, added in token order; each run's column starts as the model's
        {
            int64_t run = p.job_run[idx] - 1;
            float sw = 0.0f;
            int vcur = -1;
            T* col = nullptr;
            for (uint64_t t0 = b; t0 < e; t0 += 64) {
                const int n = (int)((e - t0) < 64 ? (e - t0) : 64);
                const uint64_t g = t0 + (uint64_t)lane;
                int wv = -1, zz = 0xFFFF;
                float tw = 1.0f;
                if (lane < n) {
                    wv = (int)p.words[g];
                    if constexpr (WT) tw = p.wt[g];
                    if ((unsigned)wv < (unsigned)p.V) {
                        const float u = uniform24(philox_x0(p.seed_lo, p.seed_hi, (uint32_t)(g - b),
                                                            (uint32_t)((g - b) >> 32), 0u, TAG_INFINIT));
                        if constexpr (ASYM) {
                            const int r = p.prior_row[wv];
                            zz = draw_from_eta_column(r >= 0 ? p.prior_eta + (int64_t)r * p.Kp : nullptr, kpl, p.eta,
                                                      p.K, u);
                        } else {
                            const float rr = u * p.alpha_cdf[p.K - 1];
                            int lo = 0, hi = p.K;
                            while (lo < hi) {
                                const int mid = (lo + hi) >> 1;
                                if (rr < p.alpha_cdf[mid]) hi = mid;
                                else lo = mid + 1;
                            }
                            zz = lo < p.K ? lo : p.K - 1;
                        }
                    }
                    p.z[g] = (uint16_t)zz;
                }
                for (int i = 0; i < n; ++i) {
                    const int v = readlane_i(wv, i);
                    if ((unsigned)v >= (unsigned)p.V) continue;
                    if (v != vcur) {
                        ++run;
                        col = cells + run * p.Kp;
                        bk_copy<CH>(col, nkwT + (int64_t)v * p.Kp, lane, kpl);
                        vcur = v;
                    }
                    const int k = readlane_i(zz, i);
                    const float w = readlane_f(tw, i);
                    bk_bump<WT>(nd, col, nk, lane, kpl, ksh, k, w, false);
                    sw += WT ? w : 1.0f;
                }
            }
            if (lane == 0) p.docw[J.slot] = sw;
        }
        for (int it = 1; it <= J.iters; ++it) {
            int64_t run = p.job_run[idx] - 1;
            int vcur = -1;
            T* col = nullptr;
            for (uint64_t t0 = b; t0 < e; t0 += 64) {
                const int n = (int)((e - t0) < 64 ? (e - t0) : 64);
                const uint64_t g = t0 + (uint64_t)lane;
                int wv = -1, zv = 0xFFFF;
                float uv = 0.0f, tw = 1.0f;
                if (lane < n) {
                    wv = (int)p.words[g];
                    zv = p.z[g];
                    if constexpr (WT) tw = p.wt[g];
                    uv = uniform24(philox_x0(p.seed_lo, p.seed_hi, (uint32_t)(g - b), (uint32_t)((g - b) >> 32),
                                             (uint32_t)it, TAG_INFER));
                }
                int zkeep = zv;
                for (int i = 0; i < n; ++i) {
                    const int v = readlane_i(wv, i);
                    if ((unsigned)v >= (unsigned)p.V) continue;
                    if (v != vcur) {
                        ++run;
                        col = cells + run * p.Kp;
                        et.word(v);
                        vcur = v;
                    }
                    const int zo = readlane_i(zv, i);
                    const int zn = bk_token<CH, WT>(nd, col, nk, p.alpha, lane, kpl, ksh, zo, readlane_f(tw, i),
                                                readlane_f(uv, i), p.K, nvalid, lastLane, et, lastc, cdf,
                                                nullptr);
                    zkeep = (lane == i) ? zn : zkeep;
                }
                if (lane < n) p.z[g] = (uint16_t)zkeep;
            }
        }
        if constexpr (ROWS) {
            bk_copy<CH>(ndg, ndl, lane, kpl);
            bk_copy<CH>(nkg, nkl, lane, kpl);
        }
        if (!p.counter) break;
    }
}

// The fold-in's log-likelihood and theta (LDAModel.hpp:931-945), in two
// passes as for K <= 2048: chunks of at most kGenChunkRuns runs add
// lgamma(cell + eta) - lgamma(model cell + eta) over the cells that changed
// (queued 64 at a time in LDS so every lane evaluates one pair), then one
// wavefront per document adds the n_dk and n_k terms and its chunk sums in
// chunk order.
template <int CH, bool WT, bool ASYM>
__global__ void __launch_bounds__(64) bk_infer_gen_runs_kernel(InferGenParams p, int kpl) {
    using T = bk_t<WT>;
    using CV = typename std::conditional<WT, fvec<CH>, ivec<CH>>::type;
    __shared__ double qa[128], qb[128];
    const int lane = lane_id();
    const T* nkw = reinterpret_cast<const T*>(p.nkw);
    const T* cells = reinterpret_cast<const T*>(p.cells);
    const int64_t cb = p.counter ? p.chunk_base[0] : p.chunk_base[p.job_base];
    const int64_t ce = p.counter ? p.chunk_base[p.njobs] : p.chunk_base[p.job_base + 1];
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int64_t ch = cb + (int64_t)blockIdx.x; ch < ce; ch += (int64_t)gridDim.x) {
        const int64_t rb = p.chunk_rb[ch], re = p.chunk_rb[ch + 1];
        double acc = 0.0;
        int qn = 0;
        for (int64_t r = rb; r < re; ++r) {
            const uint32_t v = p.run_word[r];
            int pr = -1;
            if constexpr (ASYM) pr = p.prior_row[v];
            for (int c = 0; c < kpl / CH; ++c) {
                CV m0, m1;
                if constexpr (WT) {
                    load_kvec_f<CH>(nkw + (int64_t)v * p.Kp + c * (64 * CH), lane, m0);
                    load_kvec_f<CH>(cells + r * p.Kp + c * (64 * CH), lane, m1);
                } else {
                    load_kvec<CH>(nkw + (int64_t)v * p.Kp + c * (64 * CH), lane, m0);
                    load_kvec<CH>(cells + r * p.Kp + c * (64 * CH), lane, m1);
                }
#pragma unroll
                for (int j = 0; j < CH; ++j) {
                    const T c0 = m0[j], c1 = m1[j];
                    const bool d = lane * kpl + c * CH + j < p.K && c0 != c1;
                    const unsigned long long m = __ballot(d);
                    if (!m) continue;
                    if (d) {
                        const int q = kpos_rt(kpl, lane * kpl + c * CH + j);
                        const double e = ASYM && pr >= 0 ? (double)p.prior_eta[(int64_t)pr * p.Kp + q] : p.eta_d;
                        const int at = qn + __popcll(m & below);
                        qa[at] = c1 != (T)0 ? (double)c1 + e : e;
                        qb[at] = c0 != (T)0 ? (double)c0 + e : e;
                    }
                    qn += __popcll(m);
                    if (qn >= 64) {
                        __syncthreads();
                        const int at = qn - 64 + lane;
                        acc += lgamma(qa[at]) - lgamma(qb[at]);
                        qn -= 64;
                        __syncthreads();
                    }
                }
            }
        }
        __syncthreads();
        if (lane < qn) acc += lgamma(qa[lane]) - lgamma(qb[lane]);
        acc = wave_sum_d(acc);
        if (lane == 0) p.chunk_ll[ch] = acc;
        __syncthreads();
    }
}

template <bool WT, bool ASYM>
__global__ void __launch_bounds__(64) bk_infer_gen_docs_kernel(InferGenParams p, int kpl) {
    using T = bk_t<WT>;
    const int lane = lane_id();
    for (int64_t idx = p.counter ? (int64_t)blockIdx.x : p.job_base; idx < p.njobs;
         idx += p.counter ? (int64_t)gridDim.x : p.njobs) {
        const InferJob J = p.jobs[idx];
        const T* nd = reinterpret_cast<const T*>(p.ndk_out) + (int64_t)J.slot * p.Kp;
        const T* nk1 = reinterpret_cast<const T*>(p.nk_out) + (int64_t)J.slot * p.Kp;
        const T* nk0 = reinterpret_cast<const T*>(p.nk);
        const float den = p.docw[J.slot] + p.asum_f;
        double ll = 0.0;
        for (int j = 0; j < kpl; ++j) {
            const int k = lane * kpl + j;
            if (k >= p.K) break;
            const int q = kpos_rt(kpl, k);
            const float a = p.alpha[q];
            const T x = nd[q];
            p.theta[(int64_t)J.slot * p.K + k] = ((float)x + a) / den;
            if (x != (T)0) ll += lgamma((double)x + (double)a) - lgamma((double)a);
            const double es = ASYM ? p.etasum_d[q] : p.veta_d;
            if (nk0[q] != nk1[q]) ll += lgamma((double)nk0[q] + es) - lgamma((double)nk1[q] + es);
        }
        ll = wave_sum_d(ll);
        if (lane == 0) {
            for (int64_t ch = p.chunk_base[idx]; ch < p.chunk_base[idx + 1]; ++ch) ll += p.chunk_ll[ch];
            const double asum = p.asum_d;
            ll -= lgamma((double)p.docw[J.slot] + asum) - lgamma(asum);
            p.ll[J.slot] = ll;
        }
    }
}

int bk_shift(int kpl) {
    int s = 0;
    while ((1 << s) < kpl) ++s;
    return s;
}

bool bk_kpl_ok(int kpl) { return kpl >= 1 && kpl <= kBigMaxKpl && (kpl & (kpl - 1)) == 0; }

// Dynamic LDS of the one-wavefront kernels that keep their running sums:
// Kp floats of sums and 64 chunk last values per chunk.
size_t bk_cdf_lds_bytes(int kpl, bool many = false) {
    const int ch = kpl >= 16 ? 16 : kpl;
    const size_t sums = (many && kpl > kBkCdfManyMaxKpl) ? 0 : (size_t)64 * kpl;
    return (sums + (size_t)64 * (kpl / ch)) * sizeof(float);
}

size_t bk_det_lds_bytes(int kpl) {
    return bk_cdf_lds_bytes(kpl) + (bk_det_rows(kpl) ? (size_t)2 * 64 * kpl * sizeof(int) : 0);
}

// Instantiations: CH = 16 for every count type (K above 256, which includes
// every K above 2048); CH = KPL < 16 for integer counts only (the exact-count
// path of K <= 2048, taken for documents over 2^20 tokens or alpha >= 2^20;
// float counts need no such path).
#define LDA_BK_WA(WT, ASYM, ...)                                          \
    if (WT) {                                                              \
        if (ASYM) { constexpr bool W_ = true, A_ = true; __VA_ARGS__; }           \
        else { constexpr bool W_ = true, A_ = false; __VA_ARGS__; }               \
    } else {                                                               \
        if (ASYM) { constexpr bool W_ = false, A_ = true; __VA_ARGS__; }          \
        else { constexpr bool W_ = false, A_ = false; __VA_ARGS__; }              \
    }
#define LDA_BK_A(ASYM, ...)                                               \
    if (ASYM) { constexpr bool W_ = false, A_ = true; __VA_ARGS__; }              \
    else { constexpr bool W_ = false, A_ = false; __VA_ARGS__; }
#define LDA_BK_DISPATCH(KPL_, WT, ASYM, ...)                              \
    switch ((KPL_) >= 16 ? 16 : (KPL_)) {                                  \
        case 16: { constexpr int CH_ = 16; LDA_BK_WA(WT, ASYM, __VA_ARGS__) } break; \
        case 8: if (WT) return hipErrorInvalidValue; { constexpr int CH_ = 8; LDA_BK_A(ASYM, __VA_ARGS__) } break; \
        case 4: if (WT) return hipErrorInvalidValue; { constexpr int CH_ = 4; LDA_BK_A(ASYM, __VA_ARGS__) } break; \
        case 2: if (WT) return hipErrorInvalidValue; { constexpr int CH_ = 2; LDA_BK_A(ASYM, __VA_ARGS__) } break; \
        case 1: if (WT) return hipErrorInvalidValue; { constexpr int CH_ = 1; LDA_BK_A(ASYM, __VA_ARGS__) } break; \
        default: return hipErrorInvalidValue;                              \
    }

}  // namespace

// ================================================================ launchers
hipError_t launch_bk_sweep_det(int kpl, const SweepParams& p, bool weighted, hipStream_t s) {
    if (!bk_kpl_ok(kpl)) return hipErrorInvalidValue;
    const int sh = bk_shift(kpl);
    LDA_BK_DISPATCH(kpl, weighted, p.prior_row != nullptr,
                    hipLaunchKernelGGL((bk_sweep_det_kernel<CH_, W_, A_>), dim3(p.W), dim3(64), bk_det_lds_bytes(kpl), s,
                                       p, kpl, sh));
    return hipGetLastError();
}

hipError_t launch_bk_sweep_copy(int kpl, const SweepParams& p, bool weighted, hipStream_t s) {
    if (!bk_kpl_ok(kpl)) return hipErrorInvalidValue;
    const int sh = bk_shift(kpl);
    LDA_BK_DISPATCH(kpl, weighted, p.prior_row != nullptr,
                    hipLaunchKernelGGL((bk_sweep_copy_kernel<CH_, W_, A_>), dim3(p.W), dim3(64), bk_det_lds_bytes(kpl), s,
                                       p, kpl, sh));
    return hipGetLastError();
}

size_t bk_sweep_device_lds_bytes(int kpl, int Kp, int waves, int rows) {
    const int ch = kpl >= 16 ? 16 : kpl;
    const size_t per = (size_t)(kpl / ch) * 64 + ((rows & 1) ? (size_t)Kp : 0) + ((rows & 2) ? (size_t)Kp : 0);
    return ((size_t)2 * Kp + (size_t)waves * per) * sizeof(int);
}

hipError_t launch_bk_sweep_device(int kpl, const SweepParams& p, bool weighted, int grid, int waves, hipStream_t s) {
    if (!bk_kpl_ok(kpl) || !p.bk_scratch || grid <= 0 || waves < 1 || waves > kMaxSweepWaves) return hipErrorInvalidValue;
    const int rows = p.bk_rows & 3;
    const int sh = bk_shift(kpl);
    if (waves == 1 && !rows) {
        LDA_BK_DISPATCH(kpl, weighted, p.prior_row != nullptr,
                        hipLaunchKernelGGL((bk_sweep_device_kernel<CH_, W_, A_>), dim3(grid), dim3(64), 0, s, p, kpl, sh));
    } else {
        const size_t lds = bk_sweep_device_lds_bytes(kpl, p.Kp, waves, rows);
        if (lds > 160 * 1024) return hipErrorInvalidValue;
        if (rows == 3) {
            LDA_BK_DISPATCH(kpl, weighted, p.prior_row != nullptr,
                            hipLaunchKernelGGL((bk_sweep_device_lds_kernel<CH_, W_, A_, 3>), dim3(grid),
                                               dim3(64 * waves), lds, s, p, kpl, sh));
        } else if (rows == 1) {
            LDA_BK_DISPATCH(kpl, weighted, p.prior_row != nullptr,
                            hipLaunchKernelGGL((bk_sweep_device_lds_kernel<CH_, W_, A_, 1>), dim3(grid),
                                               dim3(64 * waves), lds, s, p, kpl, sh));
        } else if (rows == 0) {
            LDA_BK_DISPATCH(kpl, weighted, p.prior_row != nullptr,
                            hipLaunchKernelGGL((bk_sweep_device_lds_kernel<CH_, W_, A_, 0>), dim3(grid),
                                               dim3(64 * waves), lds, s, p, kpl, sh));
        } else {
            return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

// Resident workgroups per CU of the LDS variant (0: it does not fit).
int bk_sweep_device_lds_blocks_per_cu(bool weighted, bool asym, int kpl, int Kp, int waves, int rows) {
    const size_t lds = bk_sweep_device_lds_bytes(kpl, Kp, waves, rows);
    if (lds > 160 * 1024) return 0;
    int n = 0;
    hipError_t e = hipSuccess;
    constexpr int CH_ = 16;
    if (rows == 3) {
        LDA_BK_WA(weighted, asym,
                  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, bk_sweep_device_lds_kernel<CH_, W_, A_, 3>,
                                                                   64 * waves, lds));
    } else if (rows == 1) {
        LDA_BK_WA(weighted, asym,
                  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, bk_sweep_device_lds_kernel<CH_, W_, A_, 1>,
                                                                   64 * waves, lds));
    } else if (rows == 0) {
        LDA_BK_WA(weighted, asym,
                  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, bk_sweep_device_lds_kernel<CH_, W_, A_, 0>,
                                                                   64 * waves, lds));
    }
    return e == hipSuccess ? n : 0;
}

// Resident one-wavefront workgroups per CU of the CH = 16 kernels (the other
// instantiations use fewer registers): the persistent sweeps take work from
// counters and the partition rounds do not wait on each other, so a count
// above the real one only queues workgroups.
int bk_sweep_blocks_per_cu(bool weighted, bool asym, bool device, int kpl) {
    int n = 0;
    hipError_t e = hipSuccess;
    constexpr int CH_ = 16;
    LDA_BK_WA(weighted, asym,
              e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
                  &n, device ? (const void*)bk_sweep_device_kernel<CH_, W_, A_> : (const void*)bk_sweep_det_kernel<CH_, W_, A_>,
                  64, device ? 0 : bk_det_lds_bytes(kpl)));
    return e == hipSuccess ? n : 0;
}

hipError_t launch_bk_sample_frozen(int kpl, const int32_t* ndk, const int32_t* nkw, const int32_t* nk,
                                   const float* alpha, float eta, float veta, int K, int Kp, int64_t n, int grid,
                                   const int64_t* doc, const uint32_t* word, const uint16_t* z_old, const float* u,
                                   uint16_t* z_new, float* total, int32_t* scratch, hipStream_t s) {
    if (!bk_kpl_ok(kpl) || !scratch || grid <= 0) return hipErrorInvalidValue;
    const int sh = bk_shift(kpl);
    LDA_BK_DISPATCH(kpl, false, false,
                    hipLaunchKernelGGL(bk_sample_frozen_kernel<CH_>, dim3(grid), dim3(64), bk_cdf_lds_bytes(kpl, true), s, ndk,
                                       nkw, nk, alpha, eta,
                                       veta, K, Kp, n, doc, word, z_old, u, z_new, total, kpl, sh, scratch));
    return hipGetLastError();
}

// The fold-in keeps a document's n_dk and n_k rows in LDS up to 128 topics
// per lane (49 KB of LDS per wavefront at 64, 66 KB at 128): 1,000 documents
// x 20 iterations at K = 4096 8.6k -> 10.0k docs/s, at K = 8192 4.6k -> 5.7k.
constexpr int kBkGenRowsMaxKpl = 128;
bool bk_infer_gen_rows(int kpl) { return kpl <= kBkGenRowsMaxKpl; }

size_t bk_infer_gen_lds_bytes(int kpl) {
    return bk_cdf_lds_bytes(kpl, true) + (bk_infer_gen_rows(kpl) ? (size_t)2 * 64 * kpl * sizeof(int) : 0);
}

hipError_t launch_bk_infer_gen(int kpl, const InferGenParams& p, int grid, bool finalize, hipStream_t s) {
    if (!bk_kpl_ok(kpl) || grid <= 0) return hipErrorInvalidValue;
    const bool wt = p.wt != nullptr, asym = p.prior_row != nullptr;
    if (finalize) {
        LDA_BK_DISPATCH(kpl, wt, asym, {
            hipLaunchKernelGGL((bk_infer_gen_runs_kernel<CH_, W_, A_>), dim3(grid), dim3(64), 0, s, p, kpl);
            hipLaunchKernelGGL((bk_infer_gen_docs_kernel<W_, A_>), dim3(grid), dim3(64), 0, s, p, kpl);
        });
    } else {
        const int sh = bk_shift(kpl);
        if (bk_infer_gen_rows(kpl)) {
            LDA_BK_DISPATCH(kpl, wt, asym,
                            hipLaunchKernelGGL((bk_infer_gen_kernel<CH_, W_, A_, true>), dim3(grid), dim3(64),
                                               bk_infer_gen_lds_bytes(kpl), s, p, kpl, sh));
        } else {
            LDA_BK_DISPATCH(kpl, wt, asym,
                            hipLaunchKernelGGL((bk_infer_gen_kernel<CH_, W_, A_, false>), dim3(grid), dim3(64),
                                               bk_infer_gen_lds_bytes(kpl), s, p, kpl, sh));
        }
    }
    return hipGetLastError();
}

int bk_infer_gen_blocks_per_cu(bool wt, bool asym, int kpl) {
    int n = 0;
    hipError_t e = hipSuccess;
    constexpr int CH_ = 16;
    if (bk_infer_gen_rows(kpl)) {
        LDA_BK_WA(wt, asym,
                  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, bk_infer_gen_kernel<CH_, W_, A_, true>, 64,
                                                                   bk_infer_gen_lds_bytes(kpl)));
    } else {
        LDA_BK_WA(wt, asym,
                  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, bk_infer_gen_kernel<CH_, W_, A_, false>, 64,
                                                                   bk_infer_gen_lds_bytes(kpl)));
    }
    return e == hipSuccess ? n : 0;
}

}  // namespace lda
