// lda_device.hip -- gfx950 kernels of the collapsed-Gibbs LDA engine.
//
// Hot path: the per-token resampling of tomoto's sampleDocument
// (src/TopicModel/LDAModel.hpp:326-354): remove the token (addWordTo<-1>,
// LDAModel.hpp:287-299), weights (n_dk+a)(n_kw+eta)/(n_k+V*eta)
// (getZLikelihoods, LDAModel.hpp:273-285), inclusive scan (sample.hpp:61-86),
// inverse-CDF search (sampleFromDiscreteAcc, sample.hpp:133-189), add the
// token back.
//
// Execution model: one 64-lane wavefront owns one document (or one
// document x vocabulary-chunk cell) and walks its tokens in order, so every
// token sees the n_dk update of the previous one, as in tomoto.  Lane l holds
// topics [l*KPL, (l+1)*KPL) of every K-vector in registers: n_dk, n_dk+alpha,
// the worker-private n_k, n_k+V*eta and its correctly rounded reciprocal,
// and the current word's n_kw column.  The column is read with coalesced
// 16-byte loads from the word-major n_kw (see kvec_pos).  Weights use the
// reciprocal with one Markstein correction, which reproduces the IEEE
// quotient; the scan is lane-local then a DPP row scan; the search is a
// ballot over lanes.  MFMA is not used: nothing here is a contraction.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "lda_internal.h"
#include "lda_kernels_common.h"

namespace lda {
namespace {

// LDS state of the sampler, every array Kp = 64*KPL entries in the lane
// layout (kpos).  Per wavefront (one document at a time):
//   A    float(n_dk) + alpha_k; n_dk itself is rint(A - alpha_k), exact while
//        n_dk + alpha_k < 2^21 (the host sends models with a document over 2^20
//        tokens or an alpha >= 2^20 to the exact-count kernels of lda_bigk.hip)
//   col  n_kw[., v] of the current word, including this wavefront's updates
// Per workgroup (the worker; its wavefronts share one private n_k):
//   nk   the worker-private n_k (LDS atomics)
//   bf   float(n_k) + V*eta and y = RN(1 / bf), rewritten after each n_k update
// A topic update is one owner-lane write per array; the per-token reads are
// conflict-free 16-byte ds_read per lane.
struct WaveLds {
    float* A;
    int* col;
    int* nk;
    float* y;
    float* bf;
};

extern __shared__ int lds_dyn[];

template <int KPL>
__device__ __forceinline__ WaveLds wave_lds() {
    constexpr int Kp = 64 * KPL;
    int* w = lds_dyn + 3 * Kp + (threadIdx.x >> 6) * (2 * Kp);
    return WaveLds{reinterpret_cast<float*>(w), w + Kp, lds_dyn, reinterpret_cast<float*>(lds_dyn + Kp),
                   reinterpret_cast<float*>(lds_dyn + 2 * Kp)};
}

// dynamic LDS bytes of a workgroup with `waves` wavefronts
template <int KPL>
constexpr size_t wave_lds_bytes(int waves = 1) { return (size_t)(3 + 2 * waves) * 64 * KPL * sizeof(int); }

template <int KPL, typename V>
__device__ __forceinline__ void lds_load(const int* a, int lane, V& o) { load_kvec<KPL>(a, lane, o); }

template <int KPL, typename V>
__device__ __forceinline__ void lds_store(int* a, int lane, const V& o) { store_kvec<KPL>(a, lane, o); }

template <int KPL, typename V>
__device__ __forceinline__ void lds_load_f(const float* a, int lane, V& o) { load_kvec_f<KPL>(a, lane, o); }

template <int KPL, typename V>
__device__ __forceinline__ void lds_store_f(float* a, int lane, const V& o) {
    store_kvec<KPL>(reinterpret_cast<int*>(a), lane, __builtin_bit_cast(ivec<KPL>, o));
}

// addWordTo<delta> for topic z on this worker's state (LDAModel.hpp:287-299):
// the owner lane updates n_dk (through A), the column and the private n_k,
// and refreshes the reciprocal.  Concurrent updates of one topic by two
// wavefronts of a workgroup keep n_k exact; y then holds the reciprocal of a
// value n_k had a moment earlier, which the Markstein step absorbs.
template <int KPL, class E>
__device__ __forceinline__ void bump(const WaveLds& S, const fvec<KPL>& al, int lane, int z, int delta, const E& et) {
    const int sl = z & (KPL - 1);
    const float a_sl = al[sl];
    if (lane == (z >> kpl_shift<KPL>())) {
        const int q = kpos<KPL>(lane, sl);
        // the n_k atomic first, then the A and column reads: all three are in
        // flight together and one wait covers them
        const int n = atomicAdd(S.nk + q, delta) + delta;
        const float a = S.A[q];
        const int c = S.col[q];
        // n_dk +/- 1 as an exact float integer; A = that + alpha is the value
        // (float)n_dk + alpha that set_ndk writes
        const float d = __builtin_rintf(a - a_sl) + (float)delta;
        S.A[q] = d + a_sl;
        S.col[q] = c + delta;
        const float b = (float)n + et.s_at(q);
        S.bf[q] = b;
        S.y[q] = recip_rn(b);
    }
}

// Document counts into A = float(n_dk) + alpha.
template <int KPL>
__device__ __forceinline__ void set_ndk(const WaveLds& S, const fvec<KPL>& al, const ivec<KPL>& nd, int lane) {
    fvec<KPL> A;
#pragma unroll
    for (int j = 0; j < KPL; ++j) A[j] = (float)nd[j] + al[j];
    lds_store_f<KPL>(S.A, lane, A);
}

// n_dk back from A.
template <int KPL>
__device__ __forceinline__ ivec<KPL> get_ndk(const WaveLds& S, const fvec<KPL>& al, int lane) {
    fvec<KPL> A;
    lds_load_f<KPL>(S.A, lane, A);
    ivec<KPL> nd;
#pragma unroll
    for (int j = 0; j < KPL; ++j) nd[j] = (int)__builtin_rintf(A[j] - al[j]);
    return nd;
}

// The eta terms of the weights (EtaHelper, LDAModel.hpp:81-117).  EtaSym:
// eta for every cell and V*eta for every topic.  EtaPrior (word priors):
// eta_kv of the current word (the word's prior column, or eta when it has
// none) and sum_v eta_kv per topic, kept per lane for the lane's topics; the
// column is reloaded when the word changes (priors are sparse: a table of
// prior columns indexed through prior_row).  Padded topics carry eta_kv = 0
// and a sum of 1, so their weight stays exactly 0.
// e(j), s(j): this lane's topic slot j (static j); s_at(q): lane-layout
// position q (the owner lane of a single-topic update).
template <int KPL>
struct EtaSym {
    float eta, veta;
    __device__ __forceinline__ float e(int) const { return eta; }
    __device__ __forceinline__ float s(int) const { return veta; }
    __device__ __forceinline__ float s_at(int) const { return veta; }
    __device__ __forceinline__ void word(const SweepParams&, int, int) {}
};

template <int KPL>
struct EtaPrior {
    fvec<KPL> ec, vs;
    const float* es;
    float eta;
    __device__ __forceinline__ float e(int j) const { return ec[j]; }
    __device__ __forceinline__ float s(int j) const { return vs[j]; }
    __device__ __forceinline__ float s_at(int q) const { return es[q]; }
    __device__ __forceinline__ void word(const SweepParams& p, int v, int lane) {
        const int r = p.prior_row[v];
        if (r >= 0) load_kvec_f<KPL>(p.prior_eta + (int64_t)r * p.Kp, lane, ec);
        else ec = eta;
    }
};

template <int KPL, bool ASYM>
struct EtaOf { using type = EtaSym<KPL>; };
template <int KPL>
struct EtaOf<KPL, true> { using type = EtaPrior<KPL>; };

template <int KPL, bool ASYM>
__device__ __forceinline__ typename EtaOf<KPL, ASYM>::type make_eta(const SweepParams& p, int lane) {
    if constexpr (ASYM) {
        EtaPrior<KPL> E;
        E.es = p.etasum;
        E.eta = p.eta;
        E.ec = p.eta;
        load_kvec_f<KPL>(p.etasum, lane, E.vs);
        return E;
    } else {
        return EtaSym<KPL>{p.eta, p.veta};
    }
}

// Private n_k from a K-vector: n_k and the reciprocals (one wavefront writes).
template <int KPL, class E>
__device__ __forceinline__ void set_nk_vec(const WaveLds& S, const ivec<KPL>& n, int lane, const E& et) {
    fvec<KPL> y, b;
#pragma unroll
    for (int j = 0; j < KPL; ++j) {
        b[j] = (float)n[j] + et.s(j);
        y[j] = recip_rn(b[j]);
    }
    lds_store<KPL>(S.nk, lane, n);
    lds_store_f<KPL>(S.y, lane, y);
    lds_store_f<KPL>(S.bf, lane, b);
}

template <int KPL, class E>
__device__ __forceinline__ void set_nk(const WaveLds& S, const int32_t* __restrict__ src, int lane, const E& et) {
    ivec<KPL> n;
    load_kvec<KPL>(src, lane, n);
    set_nk_vec<KPL>(S, n, lane, et);
}


// One token: remove z_old, weight, scan, search, add z_new.  Returns z_new
// (wave-uniform).  *total receives cdf[K-1].  Padded topics (k >= K) have
// weight exactly 0, so inside a lane the cdf of a padded slot equals that of
// the lane's last real topic: the total can be read from slot KPL-1 of lane
// (K-1)/KPL, and a lane's miss count over all slots stays below its number of
// real topics exactly when one of them is hit.
template <int KPL, class E>
__device__ __forceinline__ int token_step(const WaveLds& S, const fvec<KPL>& al, int lane, int zo, float u, int K,
                                          int nvalid, int lastLane, const E& et, float* total) {
    bump<KPL>(S, al, lane, zo, -1, et);
    fvec<KPL> A, y, bf, B;
    ivec<KPL> col;
    lds_load_f<KPL>(S.A, lane, A);
    lds_load<KPL>(S.col, lane, col);
#pragma unroll
    for (int j = 0; j < KPL; ++j) B[j] = (float)col[j] + et.e(j);
    lds_load_f<KPL>(S.bf, lane, bf);
    lds_load_f<KPL>(S.y, lane, y);
    fvec<KPL> w;
#pragma unroll
    for (int j = 0; j < KPL; ++j) {
        // ((n_dk + alpha) * (n_kw + eta)) / (n_k + V*eta); the quotient as
        // q0 = a*y, r = fma(-q0, b, a), q = fma(r, y, q0) with y = RN(1/b)
        float a = A[j] * B[j];
        float q0 = a * y[j];
        float rr = __builtin_fmaf(-q0, bf[j], a);
        w[j] = __builtin_fmaf(rr, y[j], q0);
    }
    float tot;
    const int zn = scan_search<KPL>(w, u, K, nvalid, lastLane, tot);
    bump<KPL>(S, al, lane, zn, +1, et);
    if (total) *total = tot;
    return zn;
}

// State of a worker that is one wavefront alone (the deterministic
// schedules), held in registers from token to token: A = float(n_dk) + alpha
// of the current document, the current word's column, the worker's private
// n_k, bf = float(n_k) + V*eta (sum_v eta_kv with word priors) and
// y = RN(1 / bf), each for this lane's KPL topics.  No other wavefront reads
// them, so a token touches no LDS.  The members refer to five separate
// vectors of the kernel: one vector each, indexed by a wave-uniform slot,
// stays in registers (indexed moves), where an aggregate of them would be
// placed in memory.
template <int KPL>
struct SoloState {
    fvec<KPL>& A;
    fvec<KPL>& bf;
    fvec<KPL>& y;
    ivec<KPL>& col;
    ivec<KPL>& n;
};

template <int KPL, class E>
__device__ __forceinline__ void solo_set_nk(SoloState<KPL>& st, const int32_t* __restrict__ src, int lane,
                                            const E& et) {
    load_kvec<KPL>(src, lane, st.n);
#pragma unroll
    for (int j = 0; j < KPL; ++j) {
        st.bf[j] = (float)st.n[j] + et.s(j);
        st.y[j] = recip_rn(st.bf[j]);
    }
}

// A document's counts in and out of a SoloState (set_ndk, get_ndk).
template <int KPL>
__device__ __forceinline__ void solo_set_ndk(SoloState<KPL>& st, const fvec<KPL>& al, const ivec<KPL>& nd) {
#pragma unroll
    for (int j = 0; j < KPL; ++j) st.A[j] = (float)nd[j] + al[j];
}

template <int KPL>
__device__ __forceinline__ ivec<KPL> solo_get_ndk(const SoloState<KPL>& st, const fvec<KPL>& al) {
    ivec<KPL> nd;
#pragma unroll
    for (int j = 0; j < KPL; ++j) nd[j] = (int)__builtin_rintf(st.A[j] - al[j]);
    return nd;
}

// addWordTo<delta> (LDAModel.hpp:287-299) on a SoloState: the owner lane
// updates its slot (bump's values) and stores the new cell to the word's
// n_kw row colg, whose cells the worker owns for the round.
template <int KPL, class E>
__device__ __forceinline__ void bump_solo(SoloState<KPL>& st, const fvec<KPL>& al, int lane, int z, int delta,
                                          const E& et, int32_t* colg) {
    const int sl = z & (KPL - 1);
    if (lane == (z >> kpl_shift<KPL>())) {
        const float a_sl = al[sl];
        st.A[sl] = (__builtin_rintf(st.A[sl] - a_sl) + (float)delta) + a_sl;
        const int c = st.col[sl] + delta;
        st.col[sl] = c;
        const int m = st.n[sl] + delta;
        st.n[sl] = m;
        const float b = (float)m + et.s(sl);
        st.bf[sl] = b;
        st.y[sl] = recip_rn(b);
        colg[kpos<KPL>(lane, sl)] = c;
    }
}

// token_step on a SoloState: the same operations on the same counts, hence
// the same draw.
template <int KPL, class E>
__device__ __forceinline__ int token_step_solo(SoloState<KPL>& st, const fvec<KPL>& al, int lane, int zo, float u,
                                               int K, int nvalid, int lastLane, const E& et, int32_t* colg) {
    bump_solo<KPL>(st, al, lane, zo, -1, et, colg);
    fvec<KPL> w;
#pragma unroll
    for (int j = 0; j < KPL; ++j) {
        float a = st.A[j] * ((float)st.col[j] + et.e(j));
        float q0 = a * st.y[j];
        float rr = __builtin_fmaf(-q0, st.bf[j], a);
        w[j] = __builtin_fmaf(rr, st.y[j], q0);
    }
    float tot;
    const int zn = scan_search<KPL>(w, u, K, nvalid, lastLane, tot);
    bump_solo<KPL>(st, al, lane, zn, +1, et, colg);
    return zn;
}

// n_kw[q, v] += delta with an exact atomic.  A u16 cell is updated by a 32-bit
// atomic add on the word that holds it, the delta shifted into its half:
// cells stay in [0, 2^16) (a removal takes away a token the cell counts, and
// a cell of a narrow word never exceeds the word's token count < 2^15), so
// no carry or borrow crosses into the other half.
template <bool NARROW>
__device__ __forceinline__ void col_atomic(const SweepParams& p, int v, int q, int delta) {
    if (NARROW && v >= p.nkw_hot) {
        uint32_t* w = reinterpret_cast<uint32_t*>(p.nkw16 + (int64_t)(v - p.nkw_hot) * p.Kp) + (q >> 1);
        atomicAdd(w, (uint32_t)delta << ((q & 1) * 16));
    } else {
        atomicAdd(p.nkw + (int64_t)v * p.Kp + q, delta);
    }
}

// The next word's column as loaded: int32 cells, or for a narrow word
// (NARROW, v >= nkw_hot) its u16 cells packed in pairs, unpacked only when the
// column is stored to LDS, so that the load stays in flight until the word's
// first token.
template <int KPL, bool NARROW>
__device__ __forceinline__ void fetch_col(const SweepParams& p, int v, int lane, ivec<KPL>& raw) {
    if (NARROW && v >= p.nkw_hot) {
        const uint16_t* base = p.nkw16 + (int64_t)(v - p.nkw_hot) * p.Kp;
        if constexpr (KPL >= 4) {
#pragma unroll
            for (int g = 0; g < KPL / 4; ++g) {
                const uint2 x = *reinterpret_cast<const uint2*>(base + g * 256 + lane * 4);
                raw[2 * g] = (int)x.x;
                raw[2 * g + 1] = (int)x.y;
            }
        } else if constexpr (KPL == 2) {
            raw[0] = (int)*reinterpret_cast<const uint32_t*>(base + lane * 2);
        } else {
            raw[0] = (int)base[lane];
        }
    } else {
        load_kvec<KPL>(p.nkw + (int64_t)v * p.Kp, lane, raw);
    }
}

template <int KPL, bool NARROW>
__device__ __forceinline__ void store_col(const WaveLds& S, const SweepParams& p, int v, int lane,
                                          const ivec<KPL>& raw) {
    if (NARROW && v >= p.nkw_hot) {
        ivec<KPL> o;
        if constexpr (KPL >= 4) {
#pragma unroll
            for (int g = 0; g < KPL / 4; ++g) {
                const uint32_t x = (uint32_t)raw[2 * g], y = (uint32_t)raw[2 * g + 1];
                o[4 * g] = (int)(x & 0xFFFFu);
                o[4 * g + 1] = (int)(x >> 16);
                o[4 * g + 2] = (int)(y & 0xFFFFu);
                o[4 * g + 3] = (int)(y >> 16);
            }
        } else if constexpr (KPL == 2) {
            o[0] = (int)((uint32_t)raw[0] & 0xFFFFu);
            o[1] = (int)((uint32_t)raw[0] >> 16);
        } else {
            o[0] = raw[0];
        }
        lds_store<KPL>(S.col, lane, o);
    } else {
        lds_store<KPL>(S.col, lane, raw);
    }
}

// Tokens per group of shared n_kw atomics in a device-wide sweep (the host
// sets SweepParams::col_group, a power of two up to 64; 64 otherwise).
__device__ __forceinline__ int col_group_of(const SweepParams& p) {
    const int g = p.col_group;
    return (g >= 1 && g <= 64 && (g & (g - 1)) == 0) ? g : 64;
}

// A wave-uniform 64-bit value held in scalar registers (range bounds), so
// that loops over it run on the scalar unit.
__device__ __forceinline__ uint64_t uniform_u64(uint64_t x) {
    return ((uint64_t)(uint32_t)uniform_i((int)(x >> 32)) << 32) | (uint32_t)uniform_i((int)x);
}

// Bit masks over a block of nb tokens (token t's word wv on lane t): `live`,
// the tokens that are sampled (word in range), and `pref`, the tokens after
// which the next token's word differs and is in range (its column is fetched
// one token ahead).
__device__ __forceinline__ void block_masks(int wv, int nb, int V, int lane, uint64_t& live, uint64_t& pref) {
    const int wn = __builtin_amdgcn_update_dpp(0, wv, 0x130, 0xF, 0xF, true);  // wave_shl:1, token t+1's word
    live = __ballot(lane < nb && (unsigned)wv < (unsigned)V);
    pref = __ballot(lane + 1 < nb && wn != wv && (unsigned)wn < (unsigned)V);
}

// sampleDocument over the token range [b, e) (LDAModel.hpp:326-354) for a
// device-wide worker.  Tokens come in blocks of 64, token t of a block on
// lane t (its word, z and u); the column of the next distinct word is fetched
// one token ahead and, for a narrow word, unpacked only when it is stored to
// LDS.  The n_kw changes of each group of col_group tokens go to the shared
// n_kw at the group's end as whole-wavefront atomics, lane t those of token t
// (its word, old and new topic): issued per token from single lanes they
// would hold every later load wait behind an atomic's round trip (the vector
// memory counter retires in order).  Other workers see them up to one group
// later, as they already sample from column copies; this wavefront never
// rereads a column with changes pending (a document's words come in one run
// each, and its last group is sent before the next document).
template <int KPL, class E, bool NARROW = false>
__device__ __forceinline__ void sample_range(const WaveLds& S, const fvec<KPL>& al, const SweepParams& p, uint32_t sweep,
                                             int lane, uint64_t b, uint64_t e, int nvalid, int lastLane, E& et) {
    constexpr int SH = kpl_shift<KPL>();
    const int TF = col_group_of(p);  // divides the 64-token block
    int vcur = -1, vnext = -1;
    ivec<KPL> ncol;
    b = uniform_u64(b);
    e = uniform_u64(e);
    // the current block of up to 64 tokens, token t of it on lane t, and its masks
    int nb = 0, wv = -1, zv = 0xFFFF, zkeep = 0xFFFF;
    float uv = 0.0f;
    uint64_t g = 0, live = 0, pref = 0;
    for (uint64_t t0 = b; t0 < e; t0 += TF) {
        const int off = (int)((t0 - b) & 63);
        if (off == 0) {
            nb = (int)((e - t0) < 64 ? (e - t0) : 64);
            g = t0 + (uint64_t)lane;
            wv = -1;
            zv = 0xFFFF;
            uv = 0.0f;
            if (lane < nb) {
                wv = (int)p.words[g];
                zv = p.z[g];
                uv = uniform24(philox_x0(p.seed_lo, p.seed_hi, (uint32_t)g, (uint32_t)(g >> 32), sweep, TAG_SWEEP));
            }
            block_masks(wv, nb, p.V, lane, live, pref);
            zkeep = zv;
            vnext = -1;
        }
        const int i1 = off + TF < nb ? off + TF : nb;
        for (int i = off; i < i1; ++i) {
            if (!((live >> i) & 1)) continue;
            const int v = readlane_i(wv, i);
            if (v != vcur) {
                if (v != vnext) fetch_col<KPL, NARROW>(p, v, lane, ncol);
                store_col<KPL, NARROW>(S, p, v, lane, ncol);
                et.word(p, v, lane);
                vcur = v;
            }
            if ((pref >> i) & 1) {
                const int v2 = readlane_i(wv, i + 1);
                fetch_col<KPL, NARROW>(p, v2, lane, ncol);
                vnext = v2;
            }
            const int zo = readlane_i(zv, i);
            const float u = readlane_f(uv, i);
            const int zn = token_step<KPL>(S, al, lane, zo, u, p.K, nvalid, lastLane, et, nullptr);
            zkeep = (lane == i) ? zn : zkeep;
        }
        if (lane >= off && lane < i1 && zkeep != zv && (unsigned)wv < (unsigned)p.V) {
            col_atomic<NARROW>(p, wv, kpos<KPL>(zv >> SH, zv & (KPL - 1)), -1);
            col_atomic<NARROW>(p, wv, kpos<KPL>(zkeep >> SH, zkeep & (KPL - 1)), 1);
        }
        if (i1 == nb && lane < nb) p.z[g] = (uint16_t)zkeep;
    }
}

// sampleDocument over the token range [b, e) for a one-wavefront worker of
// the deterministic schedules, its state in registers (SoloState).  The
// worker owns the cells of the words it samples for the round, so the count
// updates store them to n_kw as they go, and the current word's column
// carries every update: vcur is the word whose column st.col holds from the
// worker's previous range, and a range starting with it reads no column.
// Other columns are fetched one token ahead; the caller orders the worker's
// earlier n_kw stores before a range's reads with a fence.
template <int KPL, class E>
__device__ __forceinline__ void sample_range_solo(SoloState<KPL>& st, const fvec<KPL>& al, const SweepParams& p,
                                                  uint32_t sweep, int lane, uint64_t b, uint64_t e, int nvalid,
                                                  int lastLane, E& et, int& vcur) {
    int vnext = -1;
    int32_t* colp = vcur >= 0 ? p.nkw + (int64_t)vcur * p.Kp : nullptr;
    ivec<KPL> ncol;
    b = uniform_u64(b);
    e = uniform_u64(e);
    for (uint64_t t0 = b; t0 < e; t0 += 64) {
        const int nb = (int)((e - t0) < 64 ? (e - t0) : 64);
        const uint64_t g = t0 + (uint64_t)lane;
        int wv = -1, zv = 0xFFFF;
        float uv = 0.0f;
        if (lane < nb) {
            wv = (int)p.words[g];
            zv = p.z[g];
            uv = uniform24(philox_x0(p.seed_lo, p.seed_hi, (uint32_t)g, (uint32_t)(g >> 32), sweep, TAG_SWEEP));
        }
        uint64_t live, pref;
        block_masks(wv, nb, p.V, lane, live, pref);
        int zkeep = zv;
        vnext = -1;
        for (int i = 0; i < nb; ++i) {
            if (!((live >> i) & 1)) continue;
            const int v = readlane_i(wv, i);
            if (v != vcur) {
                colp = p.nkw + (int64_t)v * p.Kp;
                if (v != vnext) load_kvec<KPL>(colp, lane, ncol);
                st.col = ncol;
                et.word(p, v, lane);
                vcur = v;
            }
            if ((pref >> i) & 1) {
                const int v2 = readlane_i(wv, i + 1);
                load_kvec<KPL>(p.nkw + (int64_t)v2 * p.Kp, lane, ncol);
                vnext = v2;
            }
            const int zo = readlane_i(zv, i);
            const float u = readlane_f(uv, i);
            const int zn = token_step_solo<KPL>(st, al, lane, zo, u, p.K, nvalid, lastLane, et, colp);
            zkeep = (lane == i) ? zn : zkeep;
        }
        if (lane < nb) p.z[g] = (uint16_t)zkeep;
    }
}

// Device-wide sweep: every resident wavefront is a worker with a private
// n_k (snapshot nk_in at sweep start plus its own changes, as tomoto's
// partition workers keep theirs, LDAModel.hpp:484, 596-602); documents are
// taken longest-first from a global counter; n_kw is updated with atomics
// (per group of tokens, see sample_range);
// each worker adds its n_k delta to nk_out when it runs out of documents.
constexpr int kSweepWaves = kMaxSweepWaves;  // most wavefronts per workgroup of the device-wide sweep (one worker)

template <int KPL, bool ASYM, bool NARROW>
__global__ void __launch_bounds__(64 * kSweepWaves) sweep_device_kernel(SweepParams p) {
    const int lane = lane_id();
    const WaveLds S = wave_lds<KPL>();
    int nvalid, lastLane;
    lane_limits<KPL>(p.K, lane, nvalid, lastLane);
    fvec<KPL> al;
    load_kvec_f<KPL>(p.alpha, lane, al);
    auto et = make_eta<KPL, ASYM>(p, lane);
    if (threadIdx.x < 64) set_nk<KPL>(S, p.nk_in, lane, et);
    __syncthreads();
    for (;;) {
        uint32_t idx = 0;
        if (lane == 0) idx = atomicAdd(p.counter, 1u);
        idx = (uint32_t)uniform_i((int)idx);
        if ((int64_t)idx >= p.D) break;
        const uint32_t d = p.doc_order[idx];
        const uint64_t b = p.off[d], e = p.off[d + 1];
        int32_t* ndk_row = p.ndk + (int64_t)d * p.Kp;
        ivec<KPL> nd;
        load_kvec<KPL>(ndk_row, lane, nd);
        set_ndk<KPL>(S, al, nd, lane);
        sample_range<KPL, decltype(et), NARROW>(S, al, p, p.sweep, lane, b, e, nvalid, lastLane, et);
        nd = get_ndk<KPL>(S, al, lane);
        store_kvec<KPL>(ndk_row, lane, nd);
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        ivec<KPL> snap, mine;
        load_kvec<KPL>(p.nk_in, lane, snap);
        lds_load<KPL>(S.nk, lane, mine);
#pragma unroll
        for (int j = 0; j < KPL; ++j) {
            const int dlt = mine[j] - snap[j];
            if (dlt) atomicAdd(p.nk_out + kpos<KPL>(lane, j), dlt);
        }
    }
}

// One round of the deterministic partition schedule (LDAModel.hpp:373-399):
// wavefront p is worker p; it visits documents d = id*W + ((round+p) % W) in
// forShuffled order (Utils.hpp:157-178) and samples only their tokens in
// vocabulary chunk p (coff rows p, p+1).  W == 1 is the `none` scheme.
template <int KPL, bool ASYM>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2))) sweep_det_kernel(SweepParams p) {
    const int lane = lane_id();
    const int w = blockIdx.x;
    int nvalid, lastLane;
    lane_limits<KPL>(p.K, lane, nvalid, lastLane);
    fvec<KPL> al;
    load_kvec_f<KPL>(p.alpha, lane, al);
    auto et = make_eta<KPL, ASYM>(p, lane);
    int32_t* nkp = p.nk_out + (int64_t)w * p.Kp;
    fvec<KPL> sA, sbf, sy;
    ivec<KPL> scol, sn;
    SoloState<KPL> st{sA, sbf, sy, scol, sn};
    solo_set_nk<KPL>(st, nkp, lane, et);
    const int W = p.W;
    const uint32_t sweep = p.sweep_dev ? (uint32_t)uniform_i((int)*p.sweep_dev) : p.sweep;
    const int didx = (p.round + w) % W;
    const uint64_t N = (uint64_t)((p.D + (W - 1) - didx) / W);
    const uint32_t sseed = philox_x0(p.seed_lo, p.seed_hi, (uint32_t)w, (uint32_t)p.round, sweep, TAG_SHUFFLE);
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
    // Each cell's bounds and n_dk row are loaded while the previous cell is
    // sampled (a worker's documents in a round are distinct and no other
    // worker touches them), and the current word's column stays in registers
    // from one cell to the next (the chunk's cells are this worker's alone).
    auto doc_of = [&](uint64_t i) { return (int64_t)(((i + (uint64_t)sseed) * P) % N) * W + didx; };
    int vcur = -1;
    int64_t dn = 0;
    uint64_t bn = 0, en = 0;
    ivec<KPL> ndn;
    if (N) {
        dn = doc_of(0);
        bn = p.coff[(int64_t)w * p.D + dn];
        en = p.coff[(int64_t)(w + 1) * p.D + dn];
        load_kvec<KPL>(p.ndk + dn * p.Kp, lane, ndn);
    }
    for (uint64_t i = 0; i < N; ++i) {
        const int64_t d = dn;
        const uint64_t b = bn, e = en;
        const ivec<KPL> nd0 = ndn;
        // earlier cells' n_kw stores of this worker complete before this cell reads columns
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
        if (i + 1 < N) {
            dn = doc_of(i + 1);
            bn = p.coff[(int64_t)w * p.D + dn];
            en = p.coff[(int64_t)(w + 1) * p.D + dn];
            load_kvec<KPL>(p.ndk + dn * p.Kp, lane, ndn);  // (unused when the cell is empty)
        }
        if (b == e) continue;
        solo_set_ndk<KPL>(st, al, nd0);
        sample_range_solo<KPL>(st, al, p, sweep, lane, b, e, nvalid, lastLane, et, vcur);
        store_kvec<KPL>(p.ndk + d * p.Kp, lane, solo_get_ndk<KPL>(st, al));
    }
    store_kvec<KPL>(nkp, lane, st.n);
}

// copy_merge schedule (LDAModel.hpp:401-423): wavefront w is worker w with a
// private copy of n_kw (nkw + w*V*Kp) and of n_k (nk_out + w*Kp); it runs the
// tasks c = w, w+W, ... of chStride = min(8W, D) strided tasks, task c
// visiting documents id*chStride + c in forShuffled order.  Workers share no
// counts, so the result does not depend on their timing.
template <int KPL, bool ASYM>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2))) sweep_copy_kernel(SweepParams p) {
    const int lane = lane_id();
    const int w = blockIdx.x;
    int nvalid, lastLane;
    lane_limits<KPL>(p.K, lane, nvalid, lastLane);
    fvec<KPL> al;
    load_kvec_f<KPL>(p.alpha, lane, al);
    auto et = make_eta<KPL, ASYM>(p, lane);
    int32_t* nkp = p.nk_out + (int64_t)w * p.Kp;
    fvec<KPL> sA, sbf, sy;
    ivec<KPL> scol, sn;
    SoloState<KPL> st{sA, sbf, sy, scol, sn};
    solo_set_nk<KPL>(st, nkp, lane, et);
    SweepParams q = p;
    q.nkw = p.nkw + (int64_t)w * p.V * p.Kp;
    const int W = p.W;
    const int64_t chStride = (int64_t)W * 8 < p.D ? (int64_t)W * 8 : p.D;
    int vcur = -1;
    for (int64_t ch = w; ch < chStride; ch += W) {
        const uint64_t N = (uint64_t)((p.D + (chStride - 1) - ch) / chStride);
        const uint32_t sseed = philox_x0(p.seed_lo, p.seed_hi, (uint32_t)ch, 0xC0u, p.sweep, TAG_SHUFFLE);
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
        for (uint64_t i = 0; i < N; ++i) {
            const int64_t d = (int64_t)(((i + (uint64_t)sseed) * P) % N) * chStride + ch;
            const uint64_t b = p.off[d], e = p.off[d + 1];
            if (b == e) continue;
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
            int32_t* ndk_row = p.ndk + d * p.Kp;
            ivec<KPL> nd;
            load_kvec<KPL>(ndk_row, lane, nd);
            solo_set_ndk<KPL>(st, al, nd);
            // the worker's private n_kw copy: its column in registers stays current
            sample_range_solo<KPL>(st, al, q, q.sweep, lane, b, e, nvalid, lastLane, et, vcur);
            store_kvec<KPL>(ndk_row, lane, solo_get_ndk<KPL>(st, al));
        }
    }
    store_kvec<KPL>(nkp, lane, st.n);
}

// mergeState<copy_merge> (LDAModel.hpp:555-570): n_kw = old + sum_w (local_w - old).
__global__ void merge_copies_kernel(int32_t* __restrict__ nkw, const int32_t* __restrict__ loc, int64_t n, int W) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t old = nkw[i];
        int32_t acc = old;
        for (int w = 0; w < W; ++w) acc += loc[(int64_t)w * n + i] - old;
        nkw[i] = acc;
    }
}

__global__ void replicate_kernel(int32_t* __restrict__ dst, const int32_t* __restrict__ src, int64_t n, int W) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t x = src[i];
        for (int w = 0; w < W; ++w) dst[(int64_t)w * n + i] = x;
    }
}

// ------------------------------------------------ term-weighted training ----
// TermWeight IDF / PMI (LDAModel.hpp:287-299, Utils.hpp:132-155): every token
// carries a float weight w; adding it is val += w, removing it is
// val = max(val - w, 0) (tomoto's CntUpdater<true>), and the counts are float.
// The sampler state mirrors the integer one with float fields: per wavefront
// n_dk (ND) and the current word's column; per workgroup n_k and its
// reciprocal.  The weights are the same quotient ((n_dk+alpha)(n_kw+eta)) /
// (n_k+V*eta), computed with the Markstein correction.
struct WaveLdsW {
    float* nd;
    float* col;
    float* nk;
    float* y;
};

template <int KPL>
__device__ __forceinline__ WaveLdsW wave_lds_w() {
    const WaveLds S = wave_lds<KPL>();
    return WaveLdsW{S.A, reinterpret_cast<float*>(S.col), reinterpret_cast<float*>(S.nk), S.y};
}


// addWordTo<+-w> of topic z on this worker's state (owner lane).  SHARED:
// the workgroup's wavefronts share n_k (LDS atomics; y then follows a value
// n_k had a moment earlier, which the Markstein step absorbs).
template <int KPL, bool SHARED, class E>
__device__ __forceinline__ void bump_w(const WaveLdsW& S, int lane, int z, float w, bool dec, const E& et) {
    const int sl = z & (KPL - 1);
    if (lane == (z >> kpl_shift<KPL>())) {
        const int q = kpos<KPL>(lane, sl);
        // the cells' reads first, so that they are in flight together
        const float d0 = S.nd[q], c0 = S.col[q];
        float n;
        if constexpr (SHARED) {
            if (dec) atomic_sub_clamp(S.nk + q, w);
            else atomicAdd(S.nk + q, w);
            n = S.nk[q];
        } else {
            const float n0 = S.nk[q];
            n = dec ? sub_clamp(n0, w) : n0 + w;
            S.nk[q] = n;
        }
        const float c1 = dec ? sub_clamp(c0, w) : c0 + w;
        S.nd[q] = dec ? sub_clamp(d0, w) : d0 + w;
        S.col[q] = c1;
        S.y[q] = recip_rn(n + et.s_at(q));
    }
}

template <int KPL, bool SHARED, class E>
__device__ __forceinline__ int token_step_w(const WaveLdsW& S, const fvec<KPL>& al, int lane, int zo, float wt, float u,
                                            int K, int nvalid, int lastLane, const E& et) {
    bump_w<KPL, SHARED>(S, lane, zo, wt, true, et);
    fvec<KPL> nd, col, nk, y;
    lds_load_f<KPL>(S.nd, lane, nd);
    lds_load_f<KPL>(S.col, lane, col);
    lds_load_f<KPL>(S.nk, lane, nk);
    lds_load_f<KPL>(S.y, lane, y);
    fvec<KPL> w;
#pragma unroll
    for (int j = 0; j < KPL; ++j) {
        const float a = (nd[j] + al[j]) * (col[j] + et.e(j));
        const float bf = nk[j] + et.s(j);
        const float q0 = a * y[j];
        const float rr = __builtin_fmaf(-q0, bf, a);
        w[j] = __builtin_fmaf(rr, y[j], q0);
    }
    float tot;
    const int zn = scan_search<KPL>(w, u, K, nvalid, lastLane, tot);
    bump_w<KPL, SHARED>(S, lane, zn, wt, false, et);
    return zn;
}

// A one-wavefront worker's weighted state in registers (see SoloState):
// n_dk, the current word's column and the worker's n_k as floats, and
// y = RN(1 / (n_k + V*eta)), each for this lane's KPL topics.
template <int KPL>
struct SoloStateW {
    fvec<KPL>& nd;
    fvec<KPL>& col;
    fvec<KPL>& nk;
    fvec<KPL>& y;
};

// addWordTo<+-w> (bump_w's values) on a SoloStateW; the owner lane stores
// the new cell to the word's n_kw row colg, whose cells the worker owns.
template <int KPL, class E>
__device__ __forceinline__ void bump_solo_w(const SoloStateW<KPL>& st, int lane, int z, float w, bool dec,
                                            const E& et, float* colg) {
    const int sl = z & (KPL - 1);
    if (lane == (z >> kpl_shift<KPL>())) {
        const float n0 = st.nk[sl], c0 = st.col[sl], d0 = st.nd[sl];
        const float n = dec ? sub_clamp(n0, w) : n0 + w;
        const float c1 = dec ? sub_clamp(c0, w) : c0 + w;
        st.nk[sl] = n;
        st.col[sl] = c1;
        st.nd[sl] = dec ? sub_clamp(d0, w) : d0 + w;
        st.y[sl] = recip_rn(n + et.s(sl));
        colg[kpos<KPL>(lane, sl)] = c1;
    }
}

// token_step_w on a SoloStateW: the same operations, hence the same draw.
template <int KPL, class E>
__device__ __forceinline__ int token_step_solo_w(const SoloStateW<KPL>& st, const fvec<KPL>& al, int lane, int zo,
                                                 float wt, float u, int K, int nvalid, int lastLane, const E& et,
                                                 float* colg) {
    bump_solo_w<KPL>(st, lane, zo, wt, true, et, colg);
    fvec<KPL> w;
#pragma unroll
    for (int j = 0; j < KPL; ++j) {
        const float a = (st.nd[j] + al[j]) * (st.col[j] + et.e(j));
        const float bf = st.nk[j] + et.s(j);
        const float q0 = a * st.y[j];
        const float rr = __builtin_fmaf(-q0, bf, a);
        w[j] = __builtin_fmaf(rr, st.y[j], q0);
    }
    float tot;
    const int zn = scan_search<KPL>(w, u, K, nvalid, lastLane, tot);
    bump_solo_w<KPL>(st, lane, zn, wt, false, et, colg);
    return zn;
}

// sample_range_solo with token weights: sampleDocument over [b, e) for a
// one-wavefront worker of the deterministic weighted schedules (the column
// of vcur carries over from the worker's previous range).
template <int KPL, class E>
__device__ __forceinline__ void sample_range_solo_w(const SoloStateW<KPL>& st, const fvec<KPL>& al,
                                                    const SweepParams& p, uint32_t sweep, int lane, uint64_t b,
                                                    uint64_t e, int nvalid, int lastLane, E& et, int& vcur) {
    float* nkwf = reinterpret_cast<float*>(p.nkw);
    int vnext = -1;
    float* colp = vcur >= 0 ? nkwf + (int64_t)vcur * p.Kp : nullptr;
    fvec<KPL> ncol;
    b = uniform_u64(b);
    e = uniform_u64(e);
    for (uint64_t t0 = b; t0 < e; t0 += 64) {
        const int nb = (int)((e - t0) < 64 ? (e - t0) : 64);
        const uint64_t g = t0 + (uint64_t)lane;
        int wv = -1, zv = 0xFFFF;
        float uv = 0.0f, tw = 0.0f;
        if (lane < nb) {
            wv = (int)p.words[g];
            zv = p.z[g];
            tw = p.wt[g];
            uv = uniform24(philox_x0(p.seed_lo, p.seed_hi, (uint32_t)g, (uint32_t)(g >> 32), sweep, TAG_SWEEP));
        }
        uint64_t live, pref;
        block_masks(wv, nb, p.V, lane, live, pref);
        int zkeep = zv;
        vnext = -1;
        for (int i = 0; i < nb; ++i) {
            if (!((live >> i) & 1)) continue;
            const int v = readlane_i(wv, i);
            if (v != vcur) {
                colp = nkwf + (int64_t)v * p.Kp;
                if (v != vnext) load_kvec_f<KPL>(colp, lane, ncol);
                st.col = ncol;
                et.word(p, v, lane);
                vcur = v;
            }
            if ((pref >> i) & 1) {
                const int v2 = readlane_i(wv, i + 1);
                load_kvec_f<KPL>(nkwf + (int64_t)v2 * p.Kp, lane, ncol);
                vnext = v2;
            }
            const int zo = readlane_i(zv, i);
            const float w = readlane_f(tw, i);
            const float u = readlane_f(uv, i);
            const int zn = token_step_solo_w<KPL>(st, al, lane, zo, w, u, p.K, nvalid, lastLane, et, colp);
            zkeep = (lane == i) ? zn : zkeep;
        }
        if (lane < nb) p.z[g] = (uint16_t)zkeep;
    }
}

// sampleDocument over [b, e) with token weights for a device-wide worker.
// The wavefront's own state (n_dk, the column, the workgroup's n_k) follows
// tomoto's update rule exactly; the tokens' net changes go to the shared
// n_kw as float atomics, per group of tokens as in sample_range.
template <int KPL, class E>
__device__ __forceinline__ void sample_range_w(const WaveLdsW& S, const fvec<KPL>& al, const SweepParams& p,
                                               uint32_t sweep, int lane, uint64_t b, uint64_t e, int nvalid, int lastLane,
                                               E& et) {
    constexpr int SH = kpl_shift<KPL>();
    const int TF = col_group_of(p);
    float* nkwf = reinterpret_cast<float*>(p.nkw);
    int vcur = -1, vnext = -1;
    fvec<KPL> ncol;
    b = uniform_u64(b);
    e = uniform_u64(e);
    // the current block of up to 64 tokens, token t of it on lane t, and its masks
    int nb = 0, wv = -1, zv = 0xFFFF, zkeep = 0xFFFF;
    float uv = 0.0f, tw = 0.0f;
    uint64_t g = 0, live = 0, pref = 0;
    for (uint64_t t0 = b; t0 < e; t0 += TF) {
        const int off = (int)((t0 - b) & 63);
        if (off == 0) {
            nb = (int)((e - t0) < 64 ? (e - t0) : 64);
            g = t0 + (uint64_t)lane;
            wv = -1;
            zv = 0xFFFF;
            uv = 0.0f;
            tw = 0.0f;
            if (lane < nb) {
                wv = (int)p.words[g];
                zv = p.z[g];
                tw = p.wt[g];
                uv = uniform24(philox_x0(p.seed_lo, p.seed_hi, (uint32_t)g, (uint32_t)(g >> 32), sweep, TAG_SWEEP));
            }
            block_masks(wv, nb, p.V, lane, live, pref);
            zkeep = zv;
            vnext = -1;
        }
        const int i1 = off + TF < nb ? off + TF : nb;
        for (int i = off; i < i1; ++i) {
            if (!((live >> i) & 1)) continue;
            const int v = readlane_i(wv, i);
            if (v != vcur) {
                if (v != vnext) load_kvec_f<KPL>(nkwf + (int64_t)v * p.Kp, lane, ncol);
                lds_store_f<KPL>(S.col, lane, ncol);
                et.word(p, v, lane);
                vcur = v;
            }
            if ((pref >> i) & 1) {
                const int v2 = readlane_i(wv, i + 1);
                load_kvec_f<KPL>(nkwf + (int64_t)v2 * p.Kp, lane, ncol);
                vnext = v2;
            }
            const int zo = readlane_i(zv, i);
            const float w = readlane_f(tw, i);
            const float u = readlane_f(uv, i);
            const int zn = token_step_w<KPL, true>(S, al, lane, zo, w, u, p.K, nvalid, lastLane, et);
            zkeep = (lane == i) ? zn : zkeep;
        }
        // the group's changes as whole-wavefront float atomics (see
        // sample_range); a removal can leave a rounding-level negative cell,
        // which the merge's clamp max(n_kw, 0) (LDAModel.hpp:574-577) removes
        // after the sweep.  An unchanged topic adds nothing.
        if (lane >= off && lane < i1 && zkeep != zv && (unsigned)wv < (unsigned)p.V) {
            float* c = nkwf + (int64_t)wv * p.Kp;
            unsafeAtomicAdd(c + kpos<KPL>(zv >> SH, zv & (KPL - 1)), -tw);
            unsafeAtomicAdd(c + kpos<KPL>(zkeep >> SH, zkeep & (KPL - 1)), tw);
        }
        if (i1 == nb && lane < nb) p.z[g] = (uint16_t)zkeep;
    }
}

// Device-wide weighted sweep (see sweep_device_kernel).  n_k is rebuilt from
// n_kw after the sweep (launch_clamp_rowsum_f), tomoto's merge rule for
// weighted counts, so the workgroups' private n_k need no flush.
template <int KPL, bool ASYM>
__global__ void __launch_bounds__(64 * kSweepWaves) sweep_device_w_kernel(SweepParams p) {
    const int lane = lane_id();
    const WaveLdsW S = wave_lds_w<KPL>();
    int nvalid, lastLane;
    lane_limits<KPL>(p.K, lane, nvalid, lastLane);
    fvec<KPL> al;
    load_kvec_f<KPL>(p.alpha, lane, al);
    auto et = make_eta<KPL, ASYM>(p, lane);
    if (threadIdx.x < 64) {
        fvec<KPL> n, y;
        load_kvec_f<KPL>(reinterpret_cast<const float*>(p.nk_in), lane, n);
#pragma unroll
        for (int j = 0; j < KPL; ++j) y[j] = recip_rn(n[j] + et.s(j));
        lds_store_f<KPL>(S.nk, lane, n);
        lds_store_f<KPL>(S.y, lane, y);
    }
    __syncthreads();
    for (;;) {
        uint32_t idx = 0;
        if (lane == 0) idx = atomicAdd(p.counter, 1u);
        idx = (uint32_t)uniform_i((int)idx);
        if ((int64_t)idx >= p.D) break;
        const uint32_t d = p.doc_order[idx];
        const uint64_t b = p.off[d], e = p.off[d + 1];
        float* ndk_row = reinterpret_cast<float*>(p.ndk) + (int64_t)d * p.Kp;
        fvec<KPL> nd;
        load_kvec_f<KPL>(ndk_row, lane, nd);
        lds_store_f<KPL>(S.nd, lane, nd);
        sample_range_w<KPL>(S, al, p, p.sweep, lane, b, e, nvalid, lastLane, et);
        lds_load_f<KPL>(S.nd, lane, nd);
        store_kvec<KPL>(reinterpret_cast<int32_t*>(ndk_row), lane, __builtin_bit_cast(ivec<KPL>, nd));
    }
}

// One round of the deterministic weighted partition schedule (see
// sweep_det_kernel): worker-private n_k, plain stores to n_kw.
template <int KPL, bool ASYM>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2))) sweep_det_w_kernel(SweepParams p) {
    const int lane = lane_id();
    const int w = blockIdx.x;
    int nvalid, lastLane;
    lane_limits<KPL>(p.K, lane, nvalid, lastLane);
    fvec<KPL> al;
    load_kvec_f<KPL>(p.alpha, lane, al);
    auto et = make_eta<KPL, ASYM>(p, lane);
    float* nkp = reinterpret_cast<float*>(p.nk_out) + (int64_t)w * p.Kp;
    fvec<KPL> snd, scol, snk, sy;
    const SoloStateW<KPL> st{snd, scol, snk, sy};
    load_kvec_f<KPL>(nkp, lane, snk);
#pragma unroll
    for (int j = 0; j < KPL; ++j) sy[j] = recip_rn(snk[j] + et.s(j));
    int vcur = -1;
    const int W = p.W;
    const uint32_t sweep = p.sweep_dev ? (uint32_t)uniform_i((int)*p.sweep_dev) : p.sweep;
    const int didx = (p.round + w) % W;
    const uint64_t N = (uint64_t)((p.D + (W - 1) - didx) / W);
    const uint32_t sseed = philox_x0(p.seed_lo, p.seed_hi, (uint32_t)w, (uint32_t)p.round, sweep, TAG_SHUFFLE);
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
    for (uint64_t i = 0; i < N; ++i) {
        const uint64_t id = ((i + (uint64_t)sseed) * P) % N;
        const int64_t d = (int64_t)id * W + didx;
        const uint64_t b = p.coff[(int64_t)w * p.D + d], e = p.coff[(int64_t)(w + 1) * p.D + d];
        if (b == e) continue;
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
        float* ndk_row = reinterpret_cast<float*>(p.ndk) + d * p.Kp;
        load_kvec_f<KPL>(ndk_row, lane, snd);
        sample_range_solo_w<KPL>(st, al, p, sweep, lane, b, e, nvalid, lastLane, et, vcur);
        store_kvec<KPL>(reinterpret_cast<int32_t*>(ndk_row), lane, __builtin_bit_cast(ivec<KPL>, snd));
    }
    store_kvec<KPL>(reinterpret_cast<int32_t*>(nkp), lane, __builtin_bit_cast(ivec<KPL>, snk));
}

// copy_merge with token weights: the schedule of sweep_copy_kernel on
// float counts; worker w samples on its private n_kw (nkw + w*V*Kp) and n_k
// (nk_out + w*Kp) with plain stores.
template <int KPL, bool ASYM>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2))) sweep_copy_w_kernel(SweepParams p) {
    const int lane = lane_id();
    const int w = blockIdx.x;
    int nvalid, lastLane;
    lane_limits<KPL>(p.K, lane, nvalid, lastLane);
    fvec<KPL> al;
    load_kvec_f<KPL>(p.alpha, lane, al);
    auto et = make_eta<KPL, ASYM>(p, lane);
    float* nkp = reinterpret_cast<float*>(p.nk_out) + (int64_t)w * p.Kp;
    fvec<KPL> snd, scol, snk, sy;
    const SoloStateW<KPL> st{snd, scol, snk, sy};
    load_kvec_f<KPL>(nkp, lane, snk);
#pragma unroll
    for (int j = 0; j < KPL; ++j) sy[j] = recip_rn(snk[j] + et.s(j));
    int vcur = -1;
    SweepParams q = p;
    q.nkw = p.nkw + (int64_t)w * p.V * p.Kp;
    const int W = p.W;
    const int64_t chStride = (int64_t)W * 8 < p.D ? (int64_t)W * 8 : p.D;
    for (int64_t ch = w; ch < chStride; ch += W) {
        const uint64_t N = (uint64_t)((p.D + (chStride - 1) - ch) / chStride);
        const uint32_t sseed = philox_x0(p.seed_lo, p.seed_hi, (uint32_t)ch, 0xC0u, p.sweep, TAG_SHUFFLE);
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
        for (uint64_t i = 0; i < N; ++i) {
            const int64_t d = (int64_t)(((i + (uint64_t)sseed) * P) % N) * chStride + ch;
            const uint64_t b = p.off[d], e = p.off[d + 1];
            if (b == e) continue;
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
            float* ndk_row = reinterpret_cast<float*>(p.ndk) + d * p.Kp;
            load_kvec_f<KPL>(ndk_row, lane, snd);
            sample_range_solo_w<KPL>(st, al, q, q.sweep, lane, b, e, nvalid, lastLane, et, vcur);
            store_kvec<KPL>(reinterpret_cast<int32_t*>(ndk_row), lane, __builtin_bit_cast(ivec<KPL>, snd));
        }
    }
    store_kvec<KPL>(reinterpret_cast<int32_t*>(nkp), lane, __builtin_bit_cast(ivec<KPL>, snk));
}

// mergeState<copy_merge> on float counts (LDAModel.hpp:555-564):
// n_kw = local_0, then n_kw += (local_w - old) for w = 1 .. W-1 in order.
__global__ void merge_copies_f_kernel(float* __restrict__ nkw, const float* __restrict__ loc, int64_t n, int W) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float old = nkw[i];
        float acc = loc[i];
        for (int w = 1; w < W; ++w) acc = acc + (loc[(int64_t)w * n + i] - old);
        nkw[i] = acc;
    }
}

// mergeState for weighted counts (LDAModel.hpp:565-579): n_kw = max(n_kw, 0)
// cell by cell, then n_k = rowwise sum in a fixed two-level float order:
// words in blocks of kRowsumBlock, each block summed in ascending word order
// (clamp_rowsum_f_kernel, one workgroup row per block), then the block sums
// in ascending block order (rowsum_blocks_f_kernel).
__global__ void clamp_rowsum_f_kernel(float* __restrict__ nkw, int V, int Kp, float* __restrict__ part) {
    const int pos = blockIdx.y * blockDim.x + threadIdx.x;
    const int b = blockIdx.x;
    if (pos >= Kp) return;
    const int v0 = b * kRowsumBlock;
    const int v1 = min(V, v0 + kRowsumBlock);
    float s = 0.0f;
    for (int v = v0; v < v1; ++v) {
        float x = nkw[(int64_t)v * Kp + pos];
        if (x < 0.0f) {
            x = 0.0f;
            nkw[(int64_t)v * Kp + pos] = x;
        }
        s = s + x;
    }
    part[(int64_t)b * Kp + pos] = s;
}

__global__ void rowsum_blocks_f_kernel(const float* __restrict__ part, int nb, int Kp, float* __restrict__ nk) {
    const int pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= Kp) return;
    float s = 0.0f;
    for (int b = 0; b < nb; ++b) s = s + part[(int64_t)b * Kp + pos];
    nk[pos] = s;
}

// ---------------------------------------------------------------- init ----
// Initial topics z ~ Discrete(alpha) (LDAModel.hpp:754-762): r = u*cdf[K-1],
// first k with r < cdf[k] (binary search on the non-decreasing cdf).

__global__ void init_topics_kernel(const uint32_t* __restrict__ words, uint16_t* __restrict__ z, uint64_t N,
                                   const float* __restrict__ ca, int K, int V, uint32_t s0, uint32_t s1,
                                   const int32_t* __restrict__ prior_row, const float* __restrict__ prior_eta, int kpl,
                                   int Kp, float eta) {
    for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < N; g += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t v = words[g];
        if (v >= (uint32_t)V) { z[g] = 0xFFFF; continue; }
        float u = uniform24(philox_x0(s0, s1, (uint32_t)g, (uint32_t)(g >> 32), 0u, TAG_INIT));
        if (prior_row) {
            const int r = prior_row[v];
            z[g] = (uint16_t)draw_from_eta_column(r >= 0 ? prior_eta + (int64_t)r * Kp : nullptr, kpl, eta, K, u);
            continue;
        }
        float r = u * ca[K - 1];
        int lo = 0, hi = K;  // first index with r < ca[idx]
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (r < ca[mid]) hi = mid; else lo = mid + 1;
        }
        z[g] = (uint16_t)(lo < K ? lo : K - 1);
    }
}

// Counts from z, one wavefront per document (tomoto resetStatistics /
// DocumentLDA::update, LDAModel.hpp:301-315, 1163-1173).
__global__ void build_counts_kernel(const uint32_t* __restrict__ words, const uint16_t* __restrict__ z,
                                    const uint64_t* __restrict__ off, int64_t D, int32_t* ndk, int32_t* nkw,
                                    int32_t* nk, int Kp, int V, int kpl) {
    const int lane = lane_id();
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t d = wave; d < D; d += nw) {
        for (uint64_t g = off[d] + lane; g < off[d + 1]; g += 64) {
            uint32_t v = words[g];
            if (v >= (uint32_t)V) continue;
            int k = z[g];
            int pos = kpos_rt(kpl, k);
            atomicAdd(ndk + d * Kp + pos, 1);
            atomicAdd(nkw + (int64_t)v * Kp + pos, 1);
            atomicAdd(nk + pos, 1);
        }
    }
}

// Read-outs.  One topic's row of n_kw over all words (a gather of one cell
// per word column) as tomoto's _getWidsByTopic computes it
// (LDAModel.hpp:828-840): (n_kw + eta) / (n_k + V * eta) in float, the sum
// formed by the caller; C is int32_t (TermWeight ONE) or float.
template <typename C>
__global__ void topic_word_dist_kernel(const C* __restrict__ nkw, int V, int Kp, int pos, float eta, float sum,
                                       int normalize, float* __restrict__ out) {
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
        const float x = (float)nkw[(int64_t)v * Kp + pos] + eta;
        out[v] = normalize ? x / sum : x;
    }
}

// Tokens per topic over the valid tokens (tomoto's _getTopicsCount,
// LDAModel.hpp:815-826): a workgroup histogram in LDS, added to the totals.
__global__ void __launch_bounds__(256) count_topics_kernel(const uint32_t* __restrict__ words,
                                                           const uint16_t* __restrict__ z, uint64_t N, int V, int K,
                                                           unsigned long long* __restrict__ cnt) {
    extern __shared__ int lds_hist[];
    for (int k = threadIdx.x; k < K; k += blockDim.x) lds_hist[k] = 0;
    __syncthreads();
    for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < N; g += (uint64_t)gridDim.x * blockDim.x)
        if (words[g] < (uint32_t)V) atomicAdd(&lds_hist[z[g]], 1);
    __syncthreads();
    for (int k = threadIdx.x; k < K; k += blockDim.x)
        if (lds_hist[k]) atomicAdd(&cnt[k], (unsigned long long)lds_hist[k]);
}

// Streaming-bandwidth probes: every thread reads (and for the copy writes)
// 16-byte vectors in a grid-stride loop.  The read probe folds what it reads
// into a value stored only when it equals a key the data cannot produce (the
// buffer holds small integers), so the loads stay and the kernel writes
// nothing.
__global__ void __launch_bounds__(256) probe_read_kernel(const int4* __restrict__ src, int64_t n, int* __restrict__ sink) {
    int acc = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int4 x = src[i];
        acc ^= x.x ^ x.y ^ x.z ^ x.w;
    }
    if (acc == 0x7EADBEEF) sink[0] = acc;
}

__global__ void __launch_bounds__(256) probe_copy_kernel(const int4* __restrict__ src, int4* __restrict__ dst, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void pack_narrow_kernel(const int32_t* __restrict__ nkw, uint16_t* __restrict__ nkw16, int64_t off,
                                   int64_t n, uint32_t* bad) {
    bool out = false;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t x = nkw[off + i];
        out |= x < 0 || x >= (int32_t)kNarrowCf;
        nkw16[i] = (uint16_t)x;
    }
    if (out) atomicOr(bad, 1u);
}

__global__ void unpack_narrow_kernel(int32_t* __restrict__ nkw, const uint16_t* __restrict__ nkw16, int64_t off,
                                     int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        nkw[off + i] = (int32_t)nkw16[i];
}

__global__ void copy_i32_kernel(int32_t* dst, const int32_t* src, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) dst[i] = src[i];
}
__global__ void add_i32_kernel(int32_t* dst, const int32_t* src, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) dst[i] += src[i];
}
__global__ void fill_u32_kernel(uint32_t* dst, uint32_t v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) dst[i] = v;
}

// n_k = the rowwise sum of n_kw (mergeState, LDAModel.hpp:569, 578).  Integer
// sums are exact in any order, so
// blocks of `words` words (kRowsumWords, more for a vocabulary beyond 65535
// such chunks) x 256 positions add their partial sums into nk (cleared first)
// with atomics: the whole V x Kp array is read in parallel.
constexpr int kRowsumWords = 256;

__global__ void rowsum_chunks_kernel(const int32_t* __restrict__ nkw, int V, int Kp, int words, int32_t* nk) {
    const int pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= Kp) return;
    const int64_t v0 = (int64_t)blockIdx.y * words;
    const int64_t v1 = v0 + words < V ? v0 + words : V;
    int s = 0;
    for (int64_t v = v0; v < v1; ++v) s += nkw[v * Kp + pos];
    if (s) atomicAdd(nk + pos, s);
}

// ------------------------------------------------------------ loglik ------

__device__ __forceinline__ double block_sum_d(double x, double* sh) {
    x = wave_sum_d(x);
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = x;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[i];
    }
    __syncthreads();
    return t;
}

// getLLDocs + getLLRest in FP64 (LDAModel.hpp:616-666).  Block b writes its
// docs-part partial to partial[b] and its rest-part partial to partial[nblk+b].
// Word priors (LDAModel.hpp:638-650): prior_row / prior_eta give eta_kv,
// etasum_d the per-topic sum of eta_kv in double; NULL: symmetric eta.
template <typename C>
__global__ void __launch_bounds__(256) loglik_kernel(const C* __restrict__ ndk, const C* __restrict__ doclen,
                                                     int64_t D, const C* __restrict__ nkw,
                                                     const C* __restrict__ nk, const float* __restrict__ alpha,
                                                     int V, int Kp, double asum, double eta, double veta,
                                                     double* partial, const int32_t* __restrict__ prior_row,
                                                     const float* __restrict__ prior_eta,
                                                     const double* __restrict__ etasum_d) {
    __shared__ double sh[8];
    const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t nth = (int64_t)gridDim.x * blockDim.x;
    double docs = 0.0, rest = 0.0;
    const int64_t nd = D * (int64_t)Kp;
    for (int64_t i = tid; i < nd; i += nth) {
        const C c = ndk[i];
        if (c) {
            double a = (double)alpha[i % Kp];
            docs += lgamma((double)c + a) - lgamma(a);
        }
    }
    const double lga = lgamma(asum);
    for (int64_t d = tid; d < D; d += nth) docs -= lgamma((double)doclen[d] + asum) - lga;
    const double lge = lgamma(eta);
    const int64_t nv = (int64_t)V * Kp;
    for (int64_t i = tid; i < nv; i += nth) {
        const C c = nkw[i];
        if (!c) continue;
        int r = -1;
        if (prior_row) r = prior_row[i / Kp];
        if (r < 0) rest += lgamma((double)c + eta) - lge;
        else {
            const double e = (double)prior_eta[(int64_t)r * Kp + i % Kp];
            rest += lgamma((double)c + e) - lgamma(e);
        }
    }
    for (int64_t i = tid; i < Kp; i += nth) {
        // padded positions carry n_k == 0 and alpha == 0; count real topics only
        float a = alpha[i];
        const double es = etasum_d ? etasum_d[i] : veta;
        if (a > 0.0f) rest += lgamma(es) - lgamma((double)nk[i] + es);
    }
    double bd = block_sum_d(docs, sh);
    double br = block_sum_d(rest, sh);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = bd;
        partial[gridDim.x + blockIdx.x] = br;
    }
}

__global__ void doclen_kernel(const uint32_t* __restrict__ words, const uint64_t* __restrict__ off, int64_t D, int V,
                              int32_t* doclen) {
    const int lane = lane_id();
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t d = wave; d < D; d += nw) {
        int c = 0;
        for (uint64_t g = off[d] + lane; g < off[d + 1]; g += 64) c += words[g] < (uint32_t)V;
        for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o);
        if (lane == 0) doclen[d] = c;
    }
}

// digammaApprox (src/Utils/math.h:170-175), float, as tomoto evaluates it.
__device__ __forceinline__ float digamma_approx_f(float z) {
    return logf(z + 4) - 0.5f / (z + 4) - (1.0f / 12.0f) / ((z + 4) * (z + 4)) - 1.0f / z - 1.0f / (z + 1) -
           1.0f / (z + 2) - 1.0f / (z + 3);
}

// Partial sums for optimizeParameters (LDAModel.hpp:221-265): block b sums
// documents [b*D/nblk, (b+1)*D/nblk) for every topic position (threads stride
// over positions) and, in column Kp, the denominator term.
template <typename C>
__global__ void __launch_bounds__(256) alpha_sums_kernel(const C* __restrict__ ndk,
                                                         const C* __restrict__ doclen, int64_t D,
                                                         const float* __restrict__ alpha, int Kp,
                                                         const float* __restrict__ asum_p, double* part) {
    const float asum = *asum_p;
    const int64_t d0 = D * blockIdx.x / gridDim.x, d1 = D * (blockIdx.x + 1) / gridDim.x;
    for (int pos = threadIdx.x; pos <= Kp; pos += blockDim.x) {
        double s = 0.0;
        if (pos < Kp) {
            float a = alpha[pos];
            if (a > 0.0f) {
                float f0 = digamma_approx_f(a);
                int64_t zeros = 0;
                for (int64_t d = d0; d < d1; ++d) {
                    const C c = ndk[d * Kp + pos];
                    if (c) s += (double)digamma_approx_f((float)c + a);
                    else ++zeros;
                }
                s += (double)zeros * (double)f0;
            }
        } else {
            for (int64_t d = d0; d < d1; ++d) s += (double)digamma_approx_f((float)doclen[d] + asum);
        }
        part[(int64_t)blockIdx.x * (Kp + 1) + pos] = s;
    }
}

// ----------------------------------------- alpha fixed point on the device --
// optimizeParameters (LDAModel.hpp:221-265), 10 passes of
//   alpha_k <- max((S_k - D digammaT(alpha_k)) / (S - D digammaT(sum alpha)) * alpha_k, 1e-5)
// with S_k = sum_d digammaApprox(n_dk + alpha_k), S = sum_d digammaApprox(L_d + sum alpha),
// all on the device: the sums of a pass, the update and the next pass follow
// each other on the stream with no host round trip.  Integer counts group the
// terms by value (a histogram of n_dk per topic and of L_d, built once per
// call), so a pass evaluates K x max(L_d) terms instead of D x K; the same
// float terms are summed in double in another order.

// tomoto's digammaT (math.h:150-154, the LUT of LUT.hpp): the table is built
// on the host; the interpolation is the same float operations.
constexpr int kDgN = 1024, kDgM = 6400, kDgL = 4000;
__device__ float digamma_lut_dev(const float* __restrict__ pts, float x) {
    const float P = 1.0f / 1024, Q = 1.0f / 64, R = 1.0f / 4;
    if (!__builtin_isfinite(x)) return (__builtin_isnan(x) || x < 0) ? __builtin_nanf("") : __builtin_inff();
    if (x < 0) return __builtin_nanf("");
    if (x < 0.001f) {
        if (x == 0) return -__builtin_inff();
        return logf(x + 2) - 0.5f / (x + 2) - 1 / 12.f / powf(x + 2, 2) - 1 / (x + 1) - 1 / x;
    }
    if (x >= kDgN * P + kDgM * Q + (kDgL - 1) * R) return logf(x) - 0.5f / x - 1 / 12.f / powf(x, 2);
    size_t idx;
    float a, nx = x;
    if (x < kDgN * P) {
        idx = (size_t)(nx / P);
        a = (nx - idx * P) / P;
    } else {
        nx -= kDgN * P;
        if (nx < kDgM * Q) {
            idx = (size_t)(nx / Q);
            a = (nx - idx * Q) / Q;
            idx += kDgN;
        } else {
            nx -= kDgM * Q;
            idx = (size_t)(nx / R);
            a = (nx - idx * R) / R;
            idx += kDgN + kDgM;
        }
    }
    return pts[idx] + a * (pts[idx + 1] - pts[idx]);
}

// sum of alpha in topic order, in float (the order of the host's sum)
__global__ void alpha_asum_kernel(const float* __restrict__ alpha, int K, int kpl, float* out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        float s = 0.0f;
        for (int k = 0; k < K; ++k) s += alpha[kpos_rt(kpl, k)];
        *out = s;
    }
}

// one topic's update from its sum and the denominator (thread 0)
__device__ __forceinline__ void alpha_apply(float* alpha, int pos, float a, double sk, int64_t D, float denom,
                                            const float* dglut, float amax, unsigned* bad) {
    const float nom = (float)(sk - (double)D * (double)digamma_lut_dev(dglut, a));
    const float x = nom / denom * a;
    const float an = x < 1e-5f ? 1e-5f : x;  // std::max(x, 1e-5f)
    alpha[pos] = an;
    if (!(an < amax)) atomicOr(bad, 1u);
}

// direct sums: part as written by alpha_sums_kernel (nblk rows of Kp + 1);
// one workgroup per topic position
__global__ void __launch_bounds__(256) alpha_update_direct_kernel(float* __restrict__ alpha,
                                                                  const double* __restrict__ part, int nblk, int Kp,
                                                                  int64_t D, const float* __restrict__ asum_p,
                                                                  const float* __restrict__ dglut, float amax,
                                                                  unsigned* bad) {
    __shared__ double sh[8];
    const int pos = blockIdx.x;
    const float a = alpha[pos];
    if (!(a > 0.0f)) return;  // padded position (uniform over the workgroup)
    double sk = 0.0, den = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        sk += part[(int64_t)b * (Kp + 1) + pos];
        den += part[(int64_t)b * (Kp + 1) + Kp];
    }
    sk = block_sum_d(sk, sh);
    den = block_sum_d(den, sh);
    if (threadIdx.x == 0) {
        const float denom = (float)(den - (double)D * (double)digamma_lut_dev(dglut, *asum_p));
        alpha_apply(alpha, pos, a, sk, D, denom, dglut, amax, bad);
    }
}

// histograms of the integer counts: hist[pos][c] = documents with n_dk = c >= 1,
// hist_len[L] = documents of L valid tokens (every n_dk <= L <= Lmax)
__global__ void alpha_hist_kernel(const int32_t* __restrict__ ndk, const int32_t* __restrict__ doclen, int64_t D, int Kp,
                                  int Lmax, int32_t* hist, int32_t* hist_len) {
    const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t nth = (int64_t)gridDim.x * blockDim.x;
    const int64_t n = D * (int64_t)Kp;
    for (int64_t i = tid; i < n; i += nth) {
        const int32_t c = ndk[i];
        if (c > 0 && c <= Lmax) atomicAdd(hist + (i % Kp) * (int64_t)(Lmax + 1) + c, 1);
    }
    for (int64_t d = tid; d < D; d += nth) {
        const int32_t L = doclen[d];
        if (L >= 0 && L <= Lmax) atomicAdd(hist_len + L, 1);
    }
}

// S - D digammaT(sum alpha) from the length histogram; den_out[0]
__global__ void __launch_bounds__(256) alpha_den_hist_kernel(const int32_t* __restrict__ hist_len, int Lmax, int64_t D,
                                                             const float* __restrict__ asum_p,
                                                             const float* __restrict__ dglut, float* den_out) {
    __shared__ double sh[8];
    const float asum = *asum_p;
    double s = 0.0;
    for (int L = threadIdx.x; L <= Lmax; L += blockDim.x) {
        const int32_t h = hist_len[L];
        if (h) s += (double)h * (double)digamma_approx_f((float)L + asum);
    }
    s = block_sum_d(s, sh);
    if (threadIdx.x == 0) *den_out = (float)(s - (double)D * (double)digamma_lut_dev(dglut, asum));
}

__global__ void __launch_bounds__(256) alpha_update_hist_kernel(float* __restrict__ alpha, const int32_t* __restrict__ hist,
                                                                int Lmax, int64_t D, const float* __restrict__ den_p,
                                                                const float* __restrict__ dglut, float amax,
                                                                unsigned* bad) {
    __shared__ double sh[8];
    const int pos = blockIdx.x;
    const float a = alpha[pos];
    if (!(a > 0.0f)) return;  // padded position (uniform over the workgroup)
    const int32_t* h = hist + (int64_t)pos * (Lmax + 1);
    double sk = 0.0, nnz = 0.0;
    for (int c = 1 + threadIdx.x; c <= Lmax; c += blockDim.x) {
        const int32_t x = h[c];
        if (x) {
            sk += (double)x * (double)digamma_approx_f((float)c + a);
            nnz += (double)x;
        }
    }
    sk = block_sum_d(sk, sh);
    nnz = block_sum_d(nnz, sh);
    if (threadIdx.x == 0) {
        sk += (double)(D - (int64_t)nnz) * (double)digamma_approx_f(a);  // the documents with n_dk = 0
        alpha_apply(alpha, pos, a, sk, D, *den_p, dglut, amax, bad);
    }
}

// ---------------------------------------------------------- invariants ----
__global__ void recount_kernel(const uint32_t* __restrict__ words, const uint16_t* __restrict__ z,
                               const uint64_t* __restrict__ off, int64_t D, int K, int V, int Kp, int32_t* sndk,
                               int32_t* snkw, unsigned long long* flags, int kpl) {
    const int lane = lane_id();
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t d = wave; d < D; d += nw) {
        for (uint64_t g = off[d] + lane; g < off[d + 1]; g += 64) {
            uint32_t v = words[g];
            if (v >= (uint32_t)V) continue;
            int k = z[g];
            if (k >= K) { atomicAdd(flags + 6, 1ull); continue; }
            int pos = kpos_rt(kpl, k);
            atomicAdd(sndk + d * Kp + pos, 1);
            atomicAdd(snkw + (int64_t)v * Kp + pos, 1);
        }
    }
}

__global__ void compare_kernel(const int32_t* __restrict__ a, const int32_t* __restrict__ b, int64_t n,
                               unsigned long long* flag, unsigned long long* negflag) {
    unsigned long long bad = 0, neg = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        bad += a[i] != b[i];
        neg += a[i] < 0;
    }
    if (bad) atomicAdd(flag, bad);
    if (neg) atomicAdd(negflag, neg);
}

__global__ void ndk_rows_kernel(const int32_t* __restrict__ ndk, const int32_t* __restrict__ doclen, int64_t D, int Kp,
                                unsigned long long* flag) {
    const int lane = lane_id();
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t d = wave; d < D; d += nw) {
        long long s = 0;
        for (int i = lane; i < Kp; i += 64) s += ndk[d * Kp + i];
        for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
        if (lane == 0 && s != doclen[d]) atomicAdd(flag, 1ull);
    }
}

// Per topic position the sum of n_kw over the words, in 64 bits: chunks of
// `words` words add their partial sums with atomics into colsum (cleared
// first; two's complement, so negative cells of a broken state sum exactly).
__global__ void colsum_chunks_kernel(const int32_t* __restrict__ nkw, int V, int Kp, int words, long long* colsum) {
    const int pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= Kp) return;
    const int64_t v0 = (int64_t)blockIdx.y * words;
    const int64_t v1 = v0 + words < V ? v0 + words : V;
    long long s = 0;
    for (int64_t v = v0; v < v1; ++v) s += nkw[v * Kp + pos];
    if (s) atomicAdd(reinterpret_cast<unsigned long long*>(colsum + pos), (unsigned long long)s);
}

// ------------------------------------------------------ frozen sampling ----
template <int KPL>
__global__ void __launch_bounds__(64) sample_frozen_kernel(const int32_t* __restrict__ ndk, const int32_t* __restrict__ nkw,
                                                           const int32_t* __restrict__ nk, const float* __restrict__ alpha,
                                                           float eta, float veta, int K, int Kp, int64_t n,
                                                           const int64_t* __restrict__ doc, const uint32_t* __restrict__ word,
                                                           const uint16_t* __restrict__ z_old, const float* __restrict__ u,
                                                           uint16_t* z_new, float* total) {
    const int lane = lane_id();
    const WaveLds S = wave_lds<KPL>();
    int nvalid, lastLane;
    lane_limits<KPL>(K, lane, nvalid, lastLane);
    fvec<KPL> al;
    load_kvec_f<KPL>(alpha, lane, al);
    const EtaSym<KPL> et{eta, veta};
    for (int64_t t = blockIdx.x; t < n; t += gridDim.x) {
        set_nk<KPL>(S, nk, lane, et);
        ivec<KPL> x;
        load_kvec<KPL>(ndk + doc[t] * Kp, lane, x);
        set_ndk<KPL>(S, al, x, lane);
        load_kvec<KPL>(nkw + (int64_t)word[t] * Kp, lane, x);
        lds_store<KPL>(S.col, lane, x);
        float tot;
        const int zn = token_step<KPL>(S, al, lane, (int)z_old[t], u[t], K, nvalid, lastLane, et, &tot);
        if (lane == 0) {
            z_new[t] = (uint16_t)zn;
            if (total) total[t] = tot;
        }
    }
}

// -------------------------------------------------------------- fold-in ----
// Adds the current topics of the run of word v starting at token g (words are
// sorted, so the run is contiguous) into the LDS column: n_kw,eff = n_kw + c_d.
template <int KPL>
__device__ __forceinline__ void add_run_counts(const WaveLds& S, const InferParams& p, int lane, uint64_t g, uint64_t e,
                                               int v) {
    constexpr int SH = kpl_shift<KPL>();
    for (uint64_t t0 = g; t0 < e; t0 += 64) {
        const uint64_t gg = t0 + (uint64_t)lane;
        const bool in = gg < e && (int)p.words[gg] == v;
        const unsigned long long m = __ballot(in);
        const int cnt = (m == ~0ull) ? 64 : __builtin_ctzll(~m);
        const int zz = in ? (int)p.z[gg] : 0;
        for (int i = 0; i < cnt; ++i) {
            const int k = readlane_i(zz, i);
            if (lane == (k >> SH)) S.col[kpos<KPL>(lane, k & (KPL - 1))] += 1;
        }
        if (cnt < 64) break;
    }
}

// Register state of a fold-in wavefront.  The wavefront is alone with its
// document and shares no counts, so the whole sampler state stays in
// registers: six fields of KPL entries (lane l holds topics
// [l*KPL, (l+1)*KPL)):
//   A    float(n_dk) + alpha_k        NK   n_k,eff = n_k + n_dk (int bits)
//   COL  n_kw,eff = n_kw + c_d of the current word (int bits)
//   BF   float(n_k,eff) + V*eta       Y    RN(1 / BF)      AL   alpha_k
// and with word priors (ASYM) E = eta_kv of the current word and ES =
// sum_v eta_kv, which takes the place of V*eta in BF.
// A topic update indexes the state with a wave-uniform slot.  The fields are
// packed into float vectors ("banks") of 16 or 32 entries: from 16 entries
// on the compiler uses indexed register moves (s_set_gpr_idx, several under
// one index setting) rather than a select per entry, and beyond 32 it falls
// back to scratch memory.  Small KPL thus needs one bank, KPL = 32 six (eight
// with word priors).
enum FoldField { F_A = 0, F_COL = 1, F_NK = 2, F_BF = 3, F_Y = 4, F_AL = 5, F_E = 6, F_ES = 7, F_COUNT = 8 };

template <int KPL>
constexpr int fold_vw() { return F_COUNT * KPL <= 16 ? 16 : 32; }

// With at most 4 topics per lane the weight loop forms RN(1/(n_k+V*eta)) of
// every slot (3 instructions each) instead of the count update dividing once
// per bump (about 10): fewer instructions per token.
template <int KPL>
constexpr bool fold_inline_y() { return KPL <= 4; }

template <int KPL>
struct FoldBanks {
    static constexpr int VW = fold_vw<KPL>();
    static constexpr int FPV = VW / KPL;  // fields per bank
    using V = fvec<VW>;
    V &b0, &b1, &b2, &b3, &b4, &b5, &b6, &b7;

    template <int F>
    __device__ __forceinline__ V& bank() {
        constexpr int i = F / FPV;
        if constexpr (i == 0) return b0;
        else if constexpr (i == 1) return b1;
        else if constexpr (i == 2) return b2;
        else if constexpr (i == 3) return b3;
        else if constexpr (i == 4) return b4;
        else if constexpr (i == 5) return b5;
        else if constexpr (i == 6) return b6;
        else return b7;
    }
    template <int F>
    __device__ __forceinline__ float get(int sl) { return bank<F>()[(F % FPV) * KPL + sl]; }
    template <int F>
    __device__ __forceinline__ int geti(int sl) { return __float_as_int(get<F>(sl)); }
    template <int F>
    __device__ __forceinline__ void set(int sl, float x) { bank<F>()[(F % FPV) * KPL + sl] = x; }
    template <int F>
    __device__ __forceinline__ void seti(int sl, int x) { set<F>(sl, __int_as_float(x)); }
};

// addWordTo<delta> of topic z on the owner lane (LDAModel.hpp:287-299).  Every
// lane writes its slot, the non-owners their old value: a store under a lane
// condition would make the compiler copy the whole bank.
template <int KPL, bool ASYM>
__device__ __forceinline__ void fold_bump(FoldBanks<KPL>& st, int lane, int z, int delta, float veta) {
    const int sl = z & (KPL - 1);
    const bool own = lane == (z >> kpl_shift<KPL>());
    const float a = st.template get<F_AL>(sl), oA = st.template get<F_A>(sl);
    const float ob = st.template get<F_BF>(sl);
    const int oc = st.template geti<F_COL>(sl), on = st.template geti<F_NK>(sl);
    const int d = (int)__builtin_rintf(oA - a) + delta;
    const int n = on + delta;
    const float b = (float)n + (ASYM ? st.template get<F_ES>(sl) : veta);
    st.template set<F_A>(sl, own ? (float)d + a : oA);
    st.template seti<F_COL>(sl, own ? oc + delta : oc);
    st.template seti<F_NK>(sl, own ? n : on);
    st.template set<F_BF>(sl, own ? b : ob);
    if constexpr (!fold_inline_y<KPL>()) {
        const float oy = st.template get<F_Y>(sl);
        const float yn = recip_rn(b);
        st.template set<F_Y>(sl, own ? yn : oy);
    }
}

// Weight of one topic from its cells: ((n_dk + alpha) * (n_kw + eta)) /
// (n_k + V*eta), the quotient as q0 = a*y, r = fma(-q0, b, a),
// q = fma(r, y, q0) with y = RN(1/b), which is the IEEE quotient.
__device__ __forceinline__ float fold_weight(float A, int col, float bf, float y, float eta) {
    const float B = (float)col + eta;
    const float a = A * B;
    const float q0 = a * y;
    const float rr = __builtin_fmaf(-q0, bf, a);
    return __builtin_fmaf(rr, y, q0);
}

// One fold-in token: remove z_old (addWordTo<-1>), weight, scan, search, add
// z_new (addWordTo<+1>).
template <int KPL, bool ASYM>
__device__ __forceinline__ int fold_token(FoldBanks<KPL>& st, int lane, int zo, float u, int K, int nvalid,
                                          int lastLane, float eta, float veta) {
    fold_bump<KPL, ASYM>(st, lane, zo, -1, veta);
    fvec<KPL> w;
#pragma unroll
    for (int j = 0; j < KPL; ++j)
        w[j] = fold_weight(st.template get<F_A>(j), st.template geti<F_COL>(j), st.template get<F_BF>(j),
                           fold_inline_y<KPL>() ? recip_rn(st.template get<F_BF>(j)) : st.template get<F_Y>(j),
                           ASYM ? st.template get<F_E>(j) : eta);
    float tot;
    const int zn = scan_search<KPL>(w, u, K, nvalid, lastLane, tot);
    fold_bump<KPL, ASYM>(st, lane, zn, +1, veta);
    return zn;
}

// Adds one count for topic k (owner lane) to field COL.
template <int KPL>
__device__ __forceinline__ void col_add(FoldBanks<KPL>& st, int lane, int k) {
    const int sl = k & (KPL - 1);
    const int c = st.template geti<F_COL>(sl);
    st.template seti<F_COL>(sl, c + (lane == (k >> kpl_shift<KPL>()) ? 1 : 0));
}

// Adds one count for topic k (owner lane) to a KPL-entry histogram.
template <int KPL>
__device__ __forceinline__ void hist_add(ivec<fold_vw<KPL>()>& h, int lane, int k) {
    const int sl = k & (KPL - 1);
    h[sl] = h[sl] + (lane == (k >> kpl_shift<KPL>()) ? 1 : 0);
}

// c_d of word v from the tokens [g, e) of its run that lie past the current
// batch of 64 (words are sorted, so the run is contiguous).
template <int KPL>
__device__ __forceinline__ void fold_run_counts(FoldBanks<KPL>& st, const gbl<const uint32_t>* words,
                                                const gbl<uint16_t>* z, int lane, uint64_t g, uint64_t e, int v) {
    for (uint64_t t0 = g; t0 < e; t0 += 64) {
        const uint64_t gg = t0 + (uint64_t)lane;
        const bool in = gg < e && (int)words[gg] == v;
        const unsigned long long m = __ballot(in);
        const int cnt = (m == ~0ull) ? 64 : __builtin_ctzll(~m);
        const int zz = in ? (int)z[gg] : 0;
        for (int i = 0; i < cnt; ++i) col_add<KPL>(st, lane, readlane_i(zz, i));
        if (cnt < 64) break;
    }
}

// Columns of the next distinct words are fetched this many runs ahead.
template <int KPL>
constexpr int fold_prefetch() { return KPL <= 8 ? 3 : (KPL == 16 ? 2 : 1); }

// Issues the n_kw column load of the next run start in `ahead` into slot.
// The load is unconditional (an invalid word or an exhausted run list reads
// the always-present n_k vector instead): a conditional load would make the
// slot a merge of old and new values, which the compiler resolves with
// register copies that wait for the load at once.
template <int KPL>
__device__ __forceinline__ void fold_issue(ivec<KPL>& slot, unsigned long long& ahead, int wv,
                                           const gbl<const int32_t>* nkw, const gbl<const int32_t>* fallback, int V,
                                           int Kp, int lane) {
    int v = -1;
    if (ahead) {
        v = readlane_i(wv, __builtin_ctzll(ahead));
        ahead &= ahead - 1;
    }
    const gbl<const int32_t>* src = (unsigned)v < (unsigned)V ? nkw + (int64_t)v * Kp : fallback;
    gload_kvec<KPL>(src, lane, slot);
}

// Start of segment w of S of the document [b, b + n): the first word-run
// start at or after floor(n*w/S) (the same rule as the host's and the CPU
// oracle's), so that every word's tokens lie in one segment.
__device__ __forceinline__ uint64_t fold_seg_start(const gbl<const uint32_t>* words, uint64_t b, uint64_t n, int w,
                                                   int S, int lane) {
    if (w <= 0) return b;
    if (w >= S) return b + n;
    uint64_t i = (n * (uint64_t)w) / (uint64_t)S;
    if (i == 0) return b;
    for (;; i += 64) {
        const uint64_t g = i + (uint64_t)lane;
        const bool stop = g >= n || words[b + g] != words[b + g - 1];
        const unsigned long long m = __ballot(stop);
        if (m) return b + i + (uint64_t)__builtin_ctzll(m);
    }
}

// Fold-in (LDAModel.hpp:920-953) of one document per workgroup of S
// wavefronts, wavefront w sampling segment w of the document (segment rule at
// fold_segments).  The private state is kept as exact deltas: n_k,eff =
// n_k + n_dk and n_kw,eff = n_kw + c_d, c_d rebuilt per word run from the
// document's z.  With S > 1 the merged n_dk lives in LDS: every iteration
// starts from it and ends by adding each segment's change.  Every draw is
// keyed by (infer seed, token position in the document, iteration), so a
// document's result depends neither on the batch it runs in nor on how its
// iterations are split across launches.  Per batch of 64 tokens the run
// starts (tokens whose word differs from the previous one) are found with
// one ballot and their columns are loaded fold_prefetch() runs ahead into a
// register ring that the unrolled run loop indexes statically.
template <int KPL, int S, bool ASYM>
__global__ void __launch_bounds__(64 * S) infer_kernel(const InferParams* __restrict__ pp) {
    constexpr int SH = kpl_shift<KPL>();
    constexpr int P = fold_prefetch<KPL>();
    constexpr int Kp = 64 * KPL;
    constexpr int bucket = S >= 8 ? 3 : S >= 4 ? 2 : S >= 2 ? 1 : 0;
    constexpr int VW = fold_vw<KPL>();
    const InferParams p = *pp;
    const gbl<const uint32_t>* words = gptr(p.words);
    gbl<uint16_t>* zs = gptr(p.z);
    const gbl<const int32_t>* nkw = gptr(p.nkw);
    const gbl<const int32_t>* nkv = gptr(p.nk);
    const gbl<const float>* acdf = gptr(p.alpha_cdf);
    const gbl<const InferJob>* jobs = gptr(p.jobs);
    const int lane = lane_id();
    const int wave = (int)(threadIdx.x >> 6);
    int* nd_sh = lds_dyn;        // S > 1: merged n_dk (Kp entries, lane layout)
    int* job_sh = lds_dyn + Kp;  // S > 1: the workgroup's job index, then the S segment lengths
    int nvalid, lastLane;
    lane_limits<KPL>(p.K, lane, nvalid, lastLane);
    fvec<KPL> al;
    load_kvec_f<KPL>(p.alpha, lane, al);
    ivec<KPL> nk0;
    load_kvec<KPL>(p.nk, lane, nk0);
    fvec<KPL> es = 0.0f;  // word priors: sum_v eta_kv per topic (1 in padded topics)
    if constexpr (ASYM) load_kvec_f<KPL>(p.etasum, lane, es);
    const int64_t lo = p.counters ? p.bucket_lo[bucket] : 0;
    const int64_t hi = p.counters ? p.bucket_hi[bucket] : p.njobs;
    for (;;) {
        int64_t idx;
        if (p.counters) {
            if constexpr (S == 1) {
                uint32_t t = 0;
                if (lane == 0) t = atomicAdd(p.counters + bucket, 1u);
                idx = lo + (int64_t)(uint32_t)uniform_i((int)t);
            } else {
                if (threadIdx.x == 0) job_sh[0] = (int)atomicAdd(p.counters + bucket, 1u);
                __syncthreads();
                idx = lo + (int64_t)(uint32_t)uniform_i(job_sh[0]);
            }
        } else {
            idx = p.job_base;
        }
        if (idx >= hi) break;
        const uint64_t b = jobs[idx].b, e = jobs[idx].e;
        const int it_begin = jobs[idx].it_begin, it_end = jobs[idx].it_end;
        const uint64_t sb = fold_seg_start(words, b, e - b, wave, S, lane);
        const uint64_t se = fold_seg_start(words, b, e - b, wave + 1, S, lane);
        if constexpr (S > 1) {
            for (int i = threadIdx.x; i < Kp; i += 64 * S) nd_sh[i] = 0;
            if (lane == 0) job_sh[1 + wave] = (int)(se - sb);
            __syncthreads();
        }
        // batches of 64 tokens per iteration: every segment runs as many as
        // the longest one has (the merges after each batch are workgroup
        // barriers); one segment runs its own
        int nbatch = (int)((se - sb + 63) >> 6);
        if constexpr (S > 1) {
            nbatch = 0;
#pragma unroll
            for (int w = 0; w < S; ++w) nbatch = max(nbatch, (job_sh[1 + w] + 63) >> 6);
        }
        // topics at the start of this launch (initial draw or the stored z), n_dk from them
        ivec<VW> nd = 0;
        for (uint64_t t0 = sb; t0 < se; t0 += 64) {
            const uint64_t g = t0 + (uint64_t)lane;
            int zz = 0xFFFF;
            if (g < se) {
                if (it_begin == 0) {
                    const uint32_t wg = words[g];
                    if (wg < (uint32_t)p.V) {
                        const float u = uniform24(philox_x0(p.seed_lo, p.seed_hi, (uint32_t)(g - b), (uint32_t)((g - b) >> 32), 0u, TAG_INFINIT));
                        if constexpr (ASYM) {
                            // word priors: the word's eta column (LDAModel.hpp:764-775)
                            const int pr = p.prior_row[wg];
                            zz = draw_from_eta_column(pr >= 0 ? p.prior_eta + (int64_t)pr * p.Kp : nullptr, KPL, p.eta,
                                                      p.K, u);
                        } else {
                            const float r = u * acdf[p.K - 1];
                            int lo2 = 0, hi2 = p.K;
                            while (lo2 < hi2) {
                                const int mid = (lo2 + hi2) >> 1;
                                if (r < acdf[mid]) hi2 = mid; else lo2 = mid + 1;
                            }
                            zz = lo2 < p.K ? lo2 : p.K - 1;
                        }
                    }
                    zs[g] = (uint16_t)zz;
                } else {
                    zz = zs[g];
                }
            }
            const int n = (int)((se - t0) < 64 ? (se - t0) : 64);
            for (int i = 0; i < n; ++i) {
                const int k = readlane_i(zz, i);
                if (k != 0xFFFF) hist_add<KPL>(nd, lane, k);
            }
        }
        ivec<KPL> snap;  // n_dk at the start of the iteration
        if constexpr (S > 1) {
#pragma unroll
            for (int j = 0; j < KPL; ++j)
                if (nd[j]) atomicAdd(nd_sh + kpos<KPL>(lane, j), nd[j]);
        }
        fvec<VW> b0 = 0.0f, b1 = 0.0f, b2 = 0.0f, b3 = 0.0f, b4 = 0.0f, b5 = 0.0f, b6 = 0.0f, b7 = 0.0f;
        FoldBanks<KPL> st{b0, b1, b2, b3, b4, b5, b6, b7};
#pragma unroll
        for (int j = 0; j < KPL; ++j) st.template set<F_AL>(j, al[j]);
        if constexpr (ASYM) {
#pragma unroll
            for (int j = 0; j < KPL; ++j) st.template set<F_ES>(j, es[j]);
        }
        for (int it = it_begin + 1; it <= it_end; ++it) {
            int vcur = -1;
            for (int bt = 0; bt < nbatch; ++bt) {
                if constexpr (S > 1) {
                    __syncthreads();  // every segment's change is in nd_sh
                    lds_load<KPL>(nd_sh, lane, snap);
                } else if (it == it_begin + 1 && bt == 0) {
#pragma unroll
                    for (int j = 0; j < KPL; ++j) snap[j] = nd[j];
                }
                if (S > 1 || (it == it_begin + 1 && bt == 0)) {
#pragma unroll
                    for (int j = 0; j < KPL; ++j) {
                        const int n = nk0[j] + snap[j];
                        const float bb = (float)n + (ASYM ? es[j] : p.veta);
                        st.template set<F_A>(j, (float)snap[j] + al[j]);
                        st.template seti<F_NK>(j, n);
                        st.template set<F_BF>(j, bb);
                        if constexpr (!fold_inline_y<KPL>()) st.template set<F_Y>(j, 1.0f / bb);
                    }
                }
                const uint64_t t0 = sb + 64 * (uint64_t)bt;
                if (t0 < se) {
                    const int n = (int)((se - t0) < 64 ? (se - t0) : 64);
                    const uint64_t g = t0 + (uint64_t)lane;
                    int wv = -1, zv = 0xFFFF;
                    float uv = 0.0f;
                    if (lane < n) {
                        wv = (int)words[g];
                        zv = zs[g];
                        uv = uniform24(philox_x0(p.seed_lo, p.seed_hi, (uint32_t)(g - b), (uint32_t)((g - b) >> 32), (uint32_t)it, TAG_INFER));
                    }
                    // previous token's word (lane 0: the last word of the previous batch)
                    const int wp = __builtin_amdgcn_update_dpp(vcur, wv, 0x138, 0xF, 0xF, false);
                    const unsigned long long starts = __ballot(lane < n && wv != wp);
                    unsigned long long ahead = starts;
                    ivec<KPL> slot[P];
#pragma unroll
                    for (int q = 0; q < P; ++q) fold_issue<KPL>(slot[q], ahead, wv, nkw, nkv, p.V, p.Kp, lane);
                    int zkeep = zv;
                    // tokens continuing the previous batch's last run
                    const int pre = starts ? __builtin_ctzll(starts) : n;
                    if ((unsigned)vcur < (unsigned)p.V) {
                        for (int t = 0; t < pre; ++t) {
                            const int zn = fold_token<KPL, ASYM>(st, lane, readlane_i(zv, t), readlane_f(uv, t), p.K, nvalid,
                                                           lastLane, p.eta, p.veta);
                            zkeep = (lane == t) ? zn : zkeep;
                        }
                    }
                    // Runs in token order: run r takes its column from slot r % P
                    // and refills that slot with run r + P.  The unrolled q loop
                    // keeps every slot index a constant.
                    unsigned long long todo = starts;
                    while (todo) {
#pragma unroll
                        for (int q = 0; q < P; ++q) {
                            if (todo) {
                                const int i0 = __builtin_ctzll(todo);
                                todo &= todo - 1;
                                const int i1 = todo ? __builtin_ctzll(todo) : n;
                                const int v = readlane_i(wv, i0);
                                vcur = v;
#pragma unroll
                                for (int j = 0; j < KPL; ++j) st.template seti<F_COL>(j, slot[q][j]);
                                if constexpr (ASYM) {
                                    // the word's eta column: its prior row, or eta
                                    fvec<KPL> ew = p.eta;
                                    if ((unsigned)v < (unsigned)p.V) {
                                        const int pr = p.prior_row[v];
                                        if (pr >= 0) load_kvec_f<KPL>(p.prior_eta + (int64_t)pr * p.Kp, lane, ew);
                                    }
#pragma unroll
                                    for (int j = 0; j < KPL; ++j) st.template set<F_E>(j, ew[j]);
                                }
                                fold_issue<KPL>(slot[q], ahead, wv, nkw, nkv, p.V, p.Kp, lane);
                                if ((unsigned)v < (unsigned)p.V) {
                                    // c_d of this word: the run's topics from the batch
                                    // registers (not yet resampled in this iteration)
                                    // and, past the batch, from memory
                                    for (int j = i0; j < i1; ++j) col_add<KPL>(st, lane, readlane_i(zv, j));
                                    if (i1 == n) fold_run_counts<KPL>(st, words, zs, lane, t0 + (uint64_t)n, se, v);
                                    for (int t = i0; t < i1; ++t) {
                                        const int zn = fold_token<KPL, ASYM>(st, lane, readlane_i(zv, t), readlane_f(uv, t), p.K,
                                                                       nvalid, lastLane, p.eta, p.veta);
                                        zkeep = (lane == t) ? zn : zkeep;
                                    }
                                }
                            }
                        }
                    }
                    if (lane < n) zs[g] = (uint16_t)zkeep;
                }
                if constexpr (S > 1) {
                    if (it < it_end || bt + 1 < nbatch) {
                        __syncthreads();  // every segment has read nd_sh for this batch
#pragma unroll
                        for (int j = 0; j < KPL; ++j) {
                            const int dlt = (int)__builtin_rintf(st.template get<F_A>(j) - al[j]) - snap[j];
                            if (dlt) atomicAdd(nd_sh + kpos<KPL>(lane, j), dlt);
                        }
                    }
                }
            }
        }
        if constexpr (S > 1) __syncthreads();  // nd_sh and job_sh are reused by the next job
        if (!p.counters) break;
    }
}

// theta and the FP64 log-likelihood of finished fold-in jobs (it_end == iters),
// one wavefront per job: n_dk and the per-run c_d are rebuilt from z.
// getLLRest(private) - getLLRest(global) + getLLDocs(doc), with the private
// state's difference restricted to the terms the document touches
// (LDAModel.hpp:945-946); theta = (n_dk + alpha) / (L + sum alpha)
// (LDAModel.hpp:1145-1159).
template <int KPL, bool ASYM>
__global__ void __launch_bounds__(64) infer_finalize_kernel(const InferParams* __restrict__ pp) {
    constexpr int SH = kpl_shift<KPL>();
    const InferParams p = *pp;
    const int lane = lane_id();
    const WaveLds S = wave_lds<KPL>();
    int* const nd_h = S.col;                          // n_dk of the document
    int* const run_h = reinterpret_cast<int*>(S.A);   // topic counts of the open word run
    fvec<KPL> al;
    load_kvec_f<KPL>(p.alpha, lane, al);
    ivec<KPL> nk0;
    load_kvec<KPL>(p.nk, lane, nk0);
    // batched: all jobs, grid-stride; plain (no counter): the launch's own job only
    const int64_t first = p.counters ? (int64_t)blockIdx.x : p.job_base + blockIdx.x;
    const int64_t last = p.counters ? p.njobs : (first + 1 < p.njobs ? first + 1 : p.njobs);
    const int64_t stride = p.counters ? (int64_t)gridDim.x : 1;
    const double eta = (double)p.eta;
    const double veta = (double)p.V * eta;
    for (int64_t idx = first; idx < last; idx += stride) {
        const InferJob J = p.jobs[idx];
        if (J.it_end < J.iters) continue;
        const uint64_t b = J.b, e = J.e;
        ivec<KPL> zero = 0;
        lds_store<KPL>(nd_h, lane, zero);
        lds_store<KPL>(run_h, lane, zero);
        // n_dk: integer LDS atomics, order-free
        for (uint64_t t0 = b; t0 < e; t0 += 64) {
            const uint64_t g = t0 + (uint64_t)lane;
            const int zz = g < e ? (int)p.z[g] : 0xFFFF;
            if (zz != 0xFFFF) atomicAdd(nd_h + kpos<KPL>(zz >> SH, zz & (KPL - 1)), 1);
        }
        ivec<KPL> nd;
        lds_load<KPL>(nd_h, lane, nd);
        int L = 0;
#pragma unroll
        for (int j = 0; j < KPL; ++j) L += nd[j];
        for (int o = 32; o >= 1; o >>= 1) L += __shfl_xor(L, o);
        const float denom = (float)L + p.asum_f;
#pragma unroll
        for (int j = 0; j < KPL; ++j) {
            const int k = lane * KPL + j;
            if (k < p.K) p.theta[(int64_t)J.slot * p.K + k] = ((float)nd[j] + al[j]) / denom;
        }
        // getLLRest(private) - getLLRest(global) + getLLDocs(doc) over the
        // touched terms (LDAModel.hpp:616-666, 931-945): the n_k and n_dk
        // terms per topic,
        double ll = 0.0;
#pragma unroll
        for (int j = 0; j < KPL; ++j) {
            if (nd[j]) {
                const double es = ASYM ? p.etasum_d[kpos<KPL>(lane, j)] : veta;
                ll += lgamma((double)nk0[j] + es) - lgamma((double)(nk0[j] + nd[j]) + es);
                const double a = (double)al[j];
                ll += lgamma((double)nd[j] + a) - lgamma(a);
            }
        }
        // and the n_kw terms per token: a cell (v, k) holding c of the
        // document's tokens adds lgamma(n_kw + c + eta) - lgamma(n_kw + eta)
        // = sum_{r<c} log(n_kw + eta + r), so token t adds log(n_kw[v,k] +
        // eta + r_t) with r_t its rank among the document's tokens of that
        // cell.  Tokens are sorted by word, so equal cells lie in one word run;
        // r_t = (equal keys on lower lanes of the chunk) + (count of the cell
        // in earlier chunks of the same run, kept in run_h).
        int run_v = -1;
        for (uint64_t t0 = b; t0 < e; t0 += 64) {
            const int n = (int)((e - t0) < 64 ? (e - t0) : 64);
            const uint64_t g = t0 + (uint64_t)lane;
            const int v = lane < n ? (int)p.words[g] : -1;
            const bool in = lane < n && (unsigned)v < (unsigned)p.V;
            const int zz = in ? (int)p.z[g] : 0;
            int r = 0;
            for (int i = 0; i < n; ++i) {
                const int vi = readlane_i(v, i), zi = readlane_i(zz, i);
                r += (i < lane && vi == v && zi == zz) ? 1 : 0;
            }
            const int q = kpos<KPL>(zz >> SH, zz & (KPL - 1));
            if (in && v == run_v) r += run_h[q];
            if (in) {
                double e = eta;
                if constexpr (ASYM) {
                    const int pr = p.prior_row[v];
                    if (pr >= 0) e = (double)p.prior_eta[(int64_t)pr * p.Kp + q];
                }
                ll += log((double)p.nkw[(int64_t)v * p.Kp + q] + e + (double)r);
            }
            // the run still open at the end of the chunk (ids >= V sort last)
            const int vlast = readlane_i(v, n - 1);
            const int open = (unsigned)vlast < (unsigned)p.V ? vlast : -1;
            if (open != run_v) {
                lds_store<KPL>(run_h, lane, zero);
                run_v = open;
            }
            if (in && v == run_v) atomicAdd(run_h + q, 1);
        }
        ll = wave_sum_d(ll);
        if (lane == 0) {
            const double asum = p.asum_d;
            ll -= lgamma((double)L + asum) - lgamma(asum);
            p.ll[J.slot] = ll;
        }
    }
}

// ------------------------------- fold-in of term-weighted / prior models ----
// Fold-in (together=False, LDAModel.hpp:920-953) for models whose counts are
// float (TermWeight IDF / PMI) or whose eta is per word (word priors).  One
// wavefront per document, documents taken from a counter, longest first.
// The document's private copy of n_kw is kept as one column per run of equal
// word ids in `cells` (a float column's value depends on the order of its
// updates, so it cannot be rebuilt from z); n_dk, n_k and the reciprocals
// live in LDS with the training sampler's layout and update rule (bump /
// bump_w), so every weight is the training sampler's.  The initial topics
// are drawn token-parallel, then added in token order; the weight sum of the
// document is accumulated in float in the same order.
template <int KPL, bool ASYM>
__device__ __forceinline__ SweepParams gen_eta_params(const InferGenParams& p) {
    SweepParams sp{};
    sp.eta = p.eta;
    sp.veta = p.veta;
    sp.K = p.K;
    sp.V = p.V;
    sp.Kp = p.Kp;
    sp.prior_row = p.prior_row;
    sp.prior_eta = p.prior_eta;
    sp.etasum = p.etasum;
    return sp;
}

// LDS of the fold-in kernel below: per wavefront [A | col | nk | y | bf]
// (Kp each), then for S > 1 the iteration-start n_dk and n_k (Kp each).
template <int KPL>
__device__ __forceinline__ WaveLds gen_wave_lds(int wave) {
    constexpr int Kp = 64 * KPL;
    int* b = lds_dyn + wave * 5 * Kp;
    return WaveLds{reinterpret_cast<float*>(b), b + Kp, b + 2 * Kp, reinterpret_cast<float*>(b + 3 * Kp),
                   reinterpret_cast<float*>(b + 4 * Kp)};
}

template <int KPL>
constexpr size_t gen_lds_bytes(int segs) { return (size_t)(5 * segs + (segs > 1 ? 2 : 0)) * 64 * KPL * sizeof(int); }

// S wavefronts per document: S == 1 (SEG false, its own instantiation so the
// sequential sampler carries none of the segment bookkeeping) samples the
// document sequentially; S = blockDim.x / 64 > 1 (SEG) is the segmented rule (segment w = tokens
// [seg_b[w], seg_b[w+1]) of the job, its runs numbered from seg_r[w]): the
// initial topics are added by wavefront 0 in token order; every iteration
// each wavefront starts from the iteration-start n_dk / n_k, samples its
// segment, and the changes are added to the snapshot in segment order.
template <int KPL, bool WT, bool ASYM, bool SEG>
__global__ void __launch_bounds__(SEG ? 64 * kMaxFoldSegments : 64) infer_gen_kernel(InferGenParams p) {
    constexpr int SH = kpl_shift<KPL>();
    constexpr int Kp = 64 * KPL;
    const int lane = lane_id();
    // SEG: S = 2, 4 or 8 wavefronts (the workgroup size); otherwise one
    const int S = SEG ? (int)(blockDim.x >> 6) : 1;
    const int wave = SEG ? (int)(threadIdx.x >> 6) : 0;
    const WaveLds L = gen_wave_lds<KPL>(wave);
    const WaveLdsW LW{L.A, reinterpret_cast<float*>(L.col), reinterpret_cast<float*>(L.nk), L.y};
    int32_t* snap_nd = lds_dyn + S * 5 * Kp;
    int32_t* snap_nk = snap_nd + Kp;
    __shared__ int64_t job_sh;
    int nvalid, lastLane;
    lane_limits<KPL>(p.K, lane, nvalid, lastLane);
    fvec<KPL> al;
    load_kvec_f<KPL>(p.alpha, lane, al);
    const SweepParams sp = gen_eta_params<KPL, ASYM>(p);
    auto et = make_eta<KPL, ASYM>(sp, lane);
    for (;;) {
        int64_t idx;
        if (!p.counter) {
            idx = p.job_base;
        } else if (S == 1) {
            uint32_t t = 0;
            if (lane == 0) t = atomicAdd(p.counter, 1u);
            idx = (int64_t)(uint32_t)uniform_i((int)t);
        } else {
            if (threadIdx.x == 0) job_sh = (int64_t)atomicAdd(p.counter, 1u);
            __syncthreads();
            idx = job_sh;
        }
        if (idx >= p.njobs) break;
        const InferJob J = p.jobs[idx];
        const uint64_t b = J.b, e = J.e;
        if (wave == 0) {
            // private n_k = the model's, n_dk = 0
            if constexpr (WT) {
                fvec<KPL> n, y, z0 = 0.0f;
                load_kvec_f<KPL>(reinterpret_cast<const float*>(p.nk), lane, n);
#pragma unroll
                for (int j = 0; j < KPL; ++j) y[j] = recip_rn(n[j] + et.s(j));
                lds_store_f<KPL>(LW.nk, lane, n);
                lds_store_f<KPL>(LW.y, lane, y);
                lds_store_f<KPL>(LW.nd, lane, z0);
            } else {
                set_nk<KPL>(L, p.nk, lane, et);
                ivec<KPL> z0 = 0;
                set_ndk<KPL>(L, al, z0, lane);
            }
            // initial topics, added in token order; each run's column starts as the model's
            int64_t run = p.job_run[idx] - 1;
            float sw = 0.0f;
            int vcur = -1;
            int32_t* colp = nullptr;
            for (uint64_t t0 = b; t0 < e; t0 += 64) {
                const int n = (int)((e - t0) < 64 ? (e - t0) : 64);
                const uint64_t g = t0 + (uint64_t)lane;
                int wv = -1, zz = 0xFFFF;
                float tw = 1.0f;
                if (lane < n) {
                    wv = (int)p.words[g];
                    if constexpr (WT) tw = p.wt[g];
                    if ((unsigned)wv < (unsigned)p.V) {
                        const float u = uniform24(philox_x0(p.seed_lo, p.seed_hi, (uint32_t)(g - b), (uint32_t)((g - b) >> 32), 0u, TAG_INFINIT));
                        if constexpr (ASYM) {
                            const int r = p.prior_row[wv];
                            zz = draw_from_eta_column(r >= 0 ? p.prior_eta + (int64_t)r * p.Kp : nullptr, KPL, p.eta, p.K, u);
                        } else {
                            const float rr = u * p.alpha_cdf[p.K - 1];
                            int lo = 0, hi = p.K;
                            while (lo < hi) {
                                const int mid = (lo + hi) >> 1;
                                if (rr < p.alpha_cdf[mid]) hi = mid; else lo = mid + 1;
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
                        if (colp) { ivec<KPL> c; lds_load<KPL>(L.col, lane, c); store_kvec<KPL>(colp, lane, c); }
                        ++run;
                        colp = p.cells + run * p.Kp;
                        ivec<KPL> c;
                        load_kvec<KPL>(p.nkw + (int64_t)v * p.Kp, lane, c);
                        lds_store<KPL>(L.col, lane, c);
                        vcur = v;
                    }
                    const int k = readlane_i(zz, i);
                    if constexpr (WT) {
                        const float w = readlane_f(tw, i);
                        bump_w<KPL, false>(LW, lane, k, w, false, et);
                        sw += w;
                    } else {
                        bump<KPL>(L, al, lane, k, +1, et);
                        sw += 1.0f;
                    }
                }
            }
            if (colp) { ivec<KPL> c; lds_load<KPL>(L.col, lane, c); store_kvec<KPL>(colp, lane, c); }
            if (lane == 0) p.docw[J.slot] = sw;
            if (S > 1) {
                ivec<KPL> a, c;
                if constexpr (WT) {
                    fvec<KPL> x;
                    lds_load_f<KPL>(LW.nd, lane, x);
                    a = __builtin_bit_cast(ivec<KPL>, x);
                } else {
                    a = get_ndk<KPL>(L, al, lane);
                }
                lds_load<KPL>(L.nk, lane, c);
                lds_store<KPL>(snap_nd, lane, a);
                lds_store<KPL>(snap_nk, lane, c);
            }
        }
        const int64_t sbase = idx * (kMaxFoldSegments + 1);
        const uint64_t sb = S > 1 ? p.seg_b[sbase + wave] : b;
        const uint64_t se = S > 1 ? p.seg_b[sbase + wave + 1] : e;
        const int64_t run0 = S > 1 ? p.seg_r[sbase + wave] : p.job_run[idx];
        for (int it = 1; it <= J.iters; ++it) {
            if (S > 1) {
                __syncthreads();  // the snapshot (initial state or the last merge) is complete
                ivec<KPL> a, c;
                lds_load<KPL>(snap_nd, lane, a);
                lds_load<KPL>(snap_nk, lane, c);
                if constexpr (WT) {
                    const fvec<KPL> n = __builtin_bit_cast(fvec<KPL>, c);
                    fvec<KPL> y;
#pragma unroll
                    for (int j = 0; j < KPL; ++j) y[j] = recip_rn(n[j] + et.s(j));
                    lds_store<KPL>(reinterpret_cast<int*>(L.A), lane, a);
                    lds_store<KPL>(L.nk, lane, c);
                    lds_store_f<KPL>(LW.y, lane, y);
                } else {
                    set_ndk<KPL>(L, al, a, lane);
                    set_nk_vec<KPL>(L, c, lane, et);
                }
            }
            int64_t run = run0 - 1;
            int vcur = -1;
            int32_t* colp = nullptr;
            for (uint64_t t0 = sb; t0 < se; t0 += 64) {
                const int n = (int)((se - t0) < 64 ? (se - t0) : 64);
                const uint64_t g = t0 + (uint64_t)lane;
                int wv = -1, zv = 0xFFFF;
                float uv = 0.0f, tw = 1.0f;
                if (lane < n) {
                    wv = (int)p.words[g];
                    zv = p.z[g];
                    if constexpr (WT) tw = p.wt[g];
                    uv = uniform24(philox_x0(p.seed_lo, p.seed_hi, (uint32_t)(g - b), (uint32_t)((g - b) >> 32), (uint32_t)it, TAG_INFER));
                }
                int zkeep = zv;
                for (int i = 0; i < n; ++i) {
                    const int v = readlane_i(wv, i);
                    if ((unsigned)v >= (unsigned)p.V) continue;
                    if (v != vcur) {
                        ++run;
                        colp = p.cells + run * p.Kp;
                        // this lane's earlier cell stores precede the column load
                        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
                        ivec<KPL> c;
                        load_kvec<KPL>(colp, lane, c);
                        lds_store<KPL>(L.col, lane, c);
                        et.word(sp, v, lane);
                        vcur = v;
                    }
                    const int zo = readlane_i(zv, i);
                    const float u = readlane_f(uv, i);
                    int zn;
                    if constexpr (WT) {
                        zn = token_step_w<KPL, false>(LW, al, lane, zo, readlane_f(tw, i), u, p.K, nvalid, lastLane, et);
                    } else {
                        zn = token_step<KPL>(L, al, lane, zo, u, p.K, nvalid, lastLane, et, nullptr);
                    }
                    const int qo = kpos<KPL>(lane, zo & (KPL - 1)), qn = kpos<KPL>(lane, zn & (KPL - 1));
                    if (lane == (zo >> SH)) colp[qo] = L.col[qo];
                    if (lane == (zn >> SH)) colp[qn] = L.col[qn];
                    zkeep = (lane == i) ? zn : zkeep;
                }
                if (lane < n) p.z[g] = (uint16_t)zkeep;
            }
            if (S > 1) {
                __syncthreads();  // every segment is done
                // snapshot + the segments' changes, in segment order
                for (int q = (int)threadIdx.x; q < Kp; q += (int)blockDim.x) {
                    if constexpr (WT) {
                        const float a0 = __int_as_float(snap_nd[q]), c0 = __int_as_float(snap_nk[q]);
                        float a = a0, c = c0;
                        for (int w = 0; w < S; ++w) {
                            const int* wb = lds_dyn + w * 5 * Kp;
                            a = a + (__int_as_float(wb[q]) - a0);
                            c = c + (__int_as_float(wb[2 * Kp + q]) - c0);
                        }
                        snap_nd[q] = __float_as_int(a);
                        snap_nk[q] = __float_as_int(c);
                    } else {
                        const float aq = p.alpha[q];
                        const int a0 = snap_nd[q], c0 = snap_nk[q];
                        int a = a0, c = c0;
                        for (int w = 0; w < S; ++w) {
                            const int* wb = lds_dyn + w * 5 * Kp;
                            a += (int)__builtin_rintf(__int_as_float(wb[q]) - aq) - a0;
                            c += wb[2 * Kp + q] - c0;
                        }
                        snap_nd[q] = a;
                        snap_nk[q] = c;
                    }
                }
            }
        }
        // the document's n_dk and n_k for the finalize kernel
        int32_t* ndo = p.ndk_out + (int64_t)J.slot * p.Kp;
        int32_t* nko = p.nk_out + (int64_t)J.slot * p.Kp;
        if (S > 1) {
            __syncthreads();  // the last merge (or wavefront 0's initial state) is complete
            if (wave == 0) {
                ivec<KPL> a, c;
                lds_load<KPL>(snap_nd, lane, a);
                lds_load<KPL>(snap_nk, lane, c);
                store_kvec<KPL>(ndo, lane, a);
                store_kvec<KPL>(nko, lane, c);
            }
            __syncthreads();  // every wavefront is done with this job before job_sh changes
        } else if constexpr (WT) {
            fvec<KPL> a, c;
            lds_load_f<KPL>(LW.nd, lane, a);
            lds_load_f<KPL>(LW.nk, lane, c);
            store_kvec<KPL>(ndo, lane, __builtin_bit_cast(ivec<KPL>, a));
            store_kvec<KPL>(nko, lane, __builtin_bit_cast(ivec<KPL>, c));
        } else {
            store_kvec<KPL>(ndo, lane, get_ndk<KPL>(L, al, lane));
            ivec<KPL> c;
            lds_load<KPL>(L.nk, lane, c);
            store_kvec<KPL>(nko, lane, c);
        }
        if (!p.counter) break;
    }
}

// ------------------------- register-resident fold-in of IDF / PMI models ----
// The state of infer_gen_kernel<KPL, true, false> kept in registers, as the
// integer fold-in keeps its own (FoldBanks, indexed by a wave-uniform slot):
//   F_A   n_dk (float)           F_COL  the current run's private column
//   F_NK  n_k of the private copy (float)
//   F_Y   RN(1 / (F_NK + V*eta)) (formed in the weight loop for KPL <= 4)
//   F_AL  alpha_k
// The count update is bump_w's (tomoto's CntUpdater<true>: adding w,
// removing max(x - w, 0); LDAModel.hpp:287-299, Utils.hpp:132-155) and the
// weights are token_step_w's, so every draw equals the LDS kernel's.
template <int KPL>
__device__ __forceinline__ void fold_bump_w(FoldBanks<KPL>& st, int lane, int z, float w, bool dec, float veta) {
    const int sl = z & (KPL - 1);
    const bool own = lane == (z >> kpl_shift<KPL>());
    const float od = st.template get<F_A>(sl), oc = st.template get<F_COL>(sl), on = st.template get<F_NK>(sl);
    const float nd = dec ? sub_clamp(od, w) : od + w;
    const float nc = dec ? sub_clamp(oc, w) : oc + w;
    const float nn = dec ? sub_clamp(on, w) : on + w;
    st.template set<F_A>(sl, own ? nd : od);
    st.template set<F_COL>(sl, own ? nc : oc);
    st.template set<F_NK>(sl, own ? nn : on);
    if constexpr (!fold_inline_y<KPL>()) {
        const float oy = st.template get<F_Y>(sl);
        const float yn = recip_rn(nn + veta);
        st.template set<F_Y>(sl, own ? yn : oy);
    }
}

template <int KPL>
__device__ __forceinline__ int fold_token_w(FoldBanks<KPL>& st, int lane, int zo, float wt, float u, int K, int nvalid,
                                            int lastLane, float eta, float veta) {
    fold_bump_w<KPL>(st, lane, zo, wt, true, veta);
    fvec<KPL> w;
#pragma unroll
    for (int j = 0; j < KPL; ++j) {
        const float a = (st.template get<F_A>(j) + st.template get<F_AL>(j)) * (st.template get<F_COL>(j) + eta);
        const float bf = st.template get<F_NK>(j) + veta;
        const float y = fold_inline_y<KPL>() ? recip_rn(bf) : st.template get<F_Y>(j);
        const float q0 = a * y;
        const float rr = __builtin_fmaf(-q0, bf, a);
        w[j] = __builtin_fmaf(rr, y, q0);
    }
    float tot;
    const int zn = scan_search<KPL>(w, u, K, nvalid, lastLane, tot);
    fold_bump_w<KPL>(st, lane, zn, wt, false, veta);
    return zn;
}

template <int KPL>
__device__ __forceinline__ void fold_set_state_w(FoldBanks<KPL>& st, const fvec<KPL>& nd, const fvec<KPL>& nk,
                                                 float veta) {
#pragma unroll
    for (int j = 0; j < KPL; ++j) {
        st.template set<F_A>(j, nd[j]);
        st.template set<F_NK>(j, nk[j]);
        if constexpr (!fold_inline_y<KPL>()) st.template set<F_Y>(j, recip_rn(nk[j] + veta));
    }
}

template <int KPL>
__device__ __forceinline__ void fold_load_col_w(FoldBanks<KPL>& st, const fvec<KPL>& c) {
#pragma unroll
    for (int j = 0; j < KPL; ++j) st.template set<F_COL>(j, c[j]);
}

// Column loads and stores of the lanes that hold a topic below K (lanes
// from ceil(K / KPL) on hold padding only; their column entries read 0).
template <int KPL>
__device__ __forceinline__ void fold_fetch_col_w(const float* colp, int lane, bool live, fvec<KPL>& c) {
    if (live) load_kvec_f<KPL>(colp, lane, c);
    else c = fvec<KPL>(0.0f);
}

template <int KPL>
__device__ __forceinline__ void fold_store_col_w(FoldBanks<KPL>& st, float* colp, int lane, bool live) {
    fvec<KPL> c;
#pragma unroll
    for (int j = 0; j < KPL; ++j) c[j] = st.template get<F_COL>(j);
    if (live) lds_store_f<KPL>(colp, lane, c);  // store_kvec of the float bits (global memory here)
}

// The register-resident weighted fold-in holds up to 8 topics per lane (K <=
// 512) without spilling; larger K keeps the LDS-state kernel.
constexpr int kFoldWMaxKpl = 8;

// LDS of infer_w_kernel with S > 1 wavefronts: per wavefront its n_dk and n_k
// at the end of an iteration (Kp each), then the merged n_dk and n_k.
template <int KPL>
constexpr size_t gen_w_lds_bytes(int segs) { return (size_t)(segs > 1 ? 2 * segs + 2 : 0) * 64 * KPL * sizeof(float); }

// Fold-in of an IDF / PMI model without word priors: infer_gen_kernel's
// schedule (wavefront 0 adds the initial topics in token order; the
// segments start every iteration from the merged n_dk / n_k, whose changes
// are added to it in segment order) and private run columns in HBM, with the
// per-token state in registers.  The columns of a segment's runs are
// consecutive cells, so each run's column is loaded two runs ahead of its
// use (by run index) and stored when the run ends; every column is read
// once per iteration and was last stored in the previous one, so one wait
// per iteration orders the loads after those stores.
template <int KPL, bool SEG>
__global__ void __launch_bounds__(SEG ? 64 * kMaxFoldSegments : 64) infer_w_kernel(InferGenParams p) {
    constexpr int Kp = 64 * KPL;
    constexpr int VW = fold_vw<KPL>();
    const int lane = lane_id();
    const int S = SEG ? (int)(blockDim.x >> 6) : 1;
    const int wave = SEG ? (int)(threadIdx.x >> 6) : 0;
    float* lds_f = reinterpret_cast<float*>(lds_dyn);
    float* my_nd = lds_f + wave * 2 * Kp;
    float* my_nk = my_nd + Kp;
    float* snap_nd = lds_f + S * 2 * Kp;
    float* snap_nk = snap_nd + Kp;
    __shared__ int64_t job_sh;
    int nvalid, lastLane;
    lane_limits<KPL>(p.K, lane, nvalid, lastLane);
    const bool live = lane * KPL < p.K;
    fvec<KPL> al, nk0;
    load_kvec_f<KPL>(p.alpha, lane, al);
    load_kvec_f<KPL>(reinterpret_cast<const float*>(p.nk), lane, nk0);
    const float eta = p.eta, veta = p.veta;
    float* cells = reinterpret_cast<float*>(p.cells);
    const float* nkwf = reinterpret_cast<const float*>(p.nkw);
    for (;;) {
        int64_t idx;
        if (!p.counter) {
            idx = p.job_base;
        } else if (S == 1) {
            uint32_t t = 0;
            if (lane == 0) t = atomicAdd(p.counter, 1u);
            idx = (int64_t)(uint32_t)uniform_i((int)t);
        } else {
            if (threadIdx.x == 0) job_sh = (int64_t)atomicAdd(p.counter, 1u);
            __syncthreads();
            idx = job_sh;
        }
        if (idx >= p.njobs) break;
        const InferJob J = p.jobs[idx];
        const uint64_t b = J.b, e = J.e;
        fvec<VW> b0 = 0.0f, b1 = 0.0f, b2 = 0.0f, b3 = 0.0f, b4 = 0.0f, b5 = 0.0f, b6 = 0.0f, b7 = 0.0f;
        FoldBanks<KPL> st{b0, b1, b2, b3, b4, b5, b6, b7};
#pragma unroll
        for (int j = 0; j < KPL; ++j) st.template set<F_AL>(j, al[j]);
        if (wave == 0) {
            // private n_k = the model's, n_dk = 0; initial topics added in token
            // order, each run's column starting as the model's
            fold_set_state_w<KPL>(st, fvec<KPL>(0.0f), nk0, veta);
            int64_t run = p.job_run[idx] - 1;
            float sw = 0.0f;
            int vcur = -1;
            float* colp = nullptr;
            for (uint64_t t0 = b; t0 < e; t0 += 64) {
                const int n = (int)((e - t0) < 64 ? (e - t0) : 64);
                const uint64_t g = t0 + (uint64_t)lane;
                int wv = -1, zz = 0xFFFF;
                float tw = 1.0f;
                if (lane < n) {
                    wv = (int)p.words[g];
                    tw = p.wt[g];
                    if ((unsigned)wv < (unsigned)p.V) {
                        const float u = uniform24(philox_x0(p.seed_lo, p.seed_hi, (uint32_t)(g - b), (uint32_t)((g - b) >> 32), 0u, TAG_INFINIT));
                        const float rr = u * p.alpha_cdf[p.K - 1];
                        int lo = 0, hi = p.K;
                        while (lo < hi) {
                            const int mid = (lo + hi) >> 1;
                            if (rr < p.alpha_cdf[mid]) hi = mid; else lo = mid + 1;
                        }
                        zz = lo < p.K ? lo : p.K - 1;
                    }
                    p.z[g] = (uint16_t)zz;
                }
                for (int i = 0; i < n; ++i) {
                    const int v = readlane_i(wv, i);
                    if ((unsigned)v >= (unsigned)p.V) continue;
                    if (v != vcur) {
                        if (colp) fold_store_col_w<KPL>(st, colp, lane, live);
                        ++run;
                        colp = cells + run * p.Kp;
                        fvec<KPL> c;
                        fold_fetch_col_w<KPL>(nkwf + (int64_t)v * p.Kp, lane, live, c);
                        fold_load_col_w<KPL>(st, c);
                        vcur = v;
                    }
                    const float w = readlane_f(tw, i);
                    fold_bump_w<KPL>(st, lane, readlane_i(zz, i), w, false, veta);
                    sw += w;
                }
            }
            if (colp) fold_store_col_w<KPL>(st, colp, lane, live);
            if (lane == 0) p.docw[J.slot] = sw;
            if (S > 1) {
                fvec<KPL> a, c;
#pragma unroll
                for (int j = 0; j < KPL; ++j) { a[j] = st.template get<F_A>(j); c[j] = st.template get<F_NK>(j); }
                lds_store_f<KPL>(snap_nd, lane, a);
                lds_store_f<KPL>(snap_nk, lane, c);
            }
        }
        const int64_t sbase = idx * (kMaxFoldSegments + 1);
        const uint64_t sb = S > 1 ? p.seg_b[sbase + wave] : b;
        const uint64_t se = S > 1 ? p.seg_b[sbase + wave + 1] : e;
        const int64_t run0 = S > 1 ? p.seg_r[sbase + wave] : p.job_run[idx];
        const int64_t rend = S > 1 ? p.seg_r[sbase + wave + 1] : p.job_run[idx + 1];
        for (int it = 1; it <= J.iters; ++it) {
            if (S > 1) {
                __syncthreads();  // the merged state (initial or the last merge) is complete
                fvec<KPL> a, c;
                lds_load_f<KPL>(snap_nd, lane, a);
                lds_load_f<KPL>(snap_nk, lane, c);
                fold_set_state_w<KPL>(st, a, c, veta);
            } else {
                // the previous iteration's column stores precede this iteration's loads
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
            }
            int64_t run = run0 - 1;
            int vcur = -1;
            float* colp = nullptr;
            // this lane's cells of the open run changed in this pass (it owned a
            // removed or added topic): only those lanes store the column back;
            // the others' cells in HBM are still current
            bool touched = false;
            fvec<KPL> c1 = 0.0f, c2 = 0.0f;  // the columns of runs run + 1 and run + 2
            if (run0 < rend) fold_fetch_col_w<KPL>(cells + run0 * p.Kp, lane, live, c1);
            if (run0 + 1 < rend) fold_fetch_col_w<KPL>(cells + (run0 + 1) * p.Kp, lane, live, c2);
            for (uint64_t t0 = sb; t0 < se; t0 += 64) {
                const int n = (int)((se - t0) < 64 ? (se - t0) : 64);
                const uint64_t g = t0 + (uint64_t)lane;
                int wv = -1, zv = 0xFFFF;
                float uv = 0.0f, tw = 1.0f;
                if (lane < n) {
                    wv = (int)p.words[g];
                    zv = p.z[g];
                    tw = p.wt[g];
                    uv = uniform24(philox_x0(p.seed_lo, p.seed_hi, (uint32_t)(g - b), (uint32_t)((g - b) >> 32), (uint32_t)it, TAG_INFER));
                }
                int zkeep = zv;
                for (int i = 0; i < n; ++i) {
                    const int v = readlane_i(wv, i);
                    if ((unsigned)v >= (unsigned)p.V) continue;
                    if (v != vcur) {
                        if (colp) fold_store_col_w<KPL>(st, colp, lane, live && touched);
                        touched = false;
                        ++run;
                        colp = cells + run * p.Kp;
                        fold_load_col_w<KPL>(st, c1);
                        c1 = c2;
                        if (run + 2 < rend) fold_fetch_col_w<KPL>(colp + 2 * p.Kp, lane, live, c2);
                        vcur = v;
                    }
                    const int zo = readlane_i(zv, i);
                    const int zn = fold_token_w<KPL>(st, lane, zo, readlane_f(tw, i), readlane_f(uv, i), p.K, nvalid,
                                                     lastLane, eta, veta);
                    touched = touched || lane == (zo >> kpl_shift<KPL>()) || lane == (zn >> kpl_shift<KPL>());
                    zkeep = (lane == i) ? zn : zkeep;
                }
                if (lane < n) p.z[g] = (uint16_t)zkeep;
            }
            if (colp) fold_store_col_w<KPL>(st, colp, lane, live && touched);
            if (S > 1) {
                fvec<KPL> a, c;
#pragma unroll
                for (int j = 0; j < KPL; ++j) { a[j] = st.template get<F_A>(j); c[j] = st.template get<F_NK>(j); }
                lds_store_f<KPL>(my_nd, lane, a);
                lds_store_f<KPL>(my_nk, lane, c);
                __syncthreads();  // every segment is done
                // the merged state + the segments' changes, in segment order
                for (int q = (int)threadIdx.x; q < Kp; q += (int)blockDim.x) {
                    const float a0 = snap_nd[q], c0 = snap_nk[q];
                    float aa = a0, cc = c0;
                    for (int w = 0; w < S; ++w) {
                        aa = aa + (lds_f[w * 2 * Kp + q] - a0);
                        cc = cc + (lds_f[w * 2 * Kp + Kp + q] - c0);
                    }
                    snap_nd[q] = aa;
                    snap_nk[q] = cc;
                }
            }
        }
        // the document's n_dk and n_k for the finalize kernel
        float* ndo = reinterpret_cast<float*>(p.ndk_out) + (int64_t)J.slot * p.Kp;
        float* nko = reinterpret_cast<float*>(p.nk_out) + (int64_t)J.slot * p.Kp;
        if (S > 1) {
            __syncthreads();  // the last merge (or wavefront 0's initial state) is complete
            if (wave == 0) {
                fvec<KPL> a, c;
                lds_load_f<KPL>(snap_nd, lane, a);
                lds_load_f<KPL>(snap_nk, lane, c);
                lds_store_f<KPL>(ndo, lane, a);
                lds_store_f<KPL>(nko, lane, c);
            }
            __syncthreads();  // every wavefront is done with this job before job_sh and the LDS change
        } else {
            fvec<KPL> a, c;
#pragma unroll
            for (int j = 0; j < KPL; ++j) { a[j] = st.template get<F_A>(j); c[j] = st.template get<F_NK>(j); }
            lds_store_f<KPL>(ndo, lane, a);
            lds_store_f<KPL>(nko, lane, c);
        }
        if (!p.counter) break;
    }
}

// theta = (n_dk + alpha) / (weight sum + sum alpha) and the FP64
// log-likelihood getLLRest(private) - getLLRest(model) + getLLDocs(doc)
// (LDAModel.hpp:616-674, 945-946) in two launches.  The run pass sums the
// n_kw terms of the cells that changed, gen_cell_term(private) -
// gen_cell_term(model) = lgamma(a) - lgamma(b) with a = private + e (e when
// the private cell is 0: tomoto skips zero cells, and lgamma(0 + e) -
// lgamma(e) is 0 either way) and b likewise from the model's cell, over
// chunks of at most kGenChunkRuns runs of one document, one wavefront per
// chunk: the (a, b) pairs of the changed cells queue in LDS and are taken 64
// at a time, so every lane evaluates lgamma however few of its own cells
// changed.  The document pass adds the n_k terms of the topics whose n_k
// changed, the document terms and the document's chunk sums in chunk
// order; one wavefront per document.  Both sums are fixed by the document
// alone, so a document's ll does not depend on the batch around it.
template <int KPL, bool WT, bool ASYM>
__global__ void __launch_bounds__(64) infer_gen_runs_kernel(InferGenParams p) {
    using C = typename std::conditional<WT, float, int32_t>::type;
    using CV = typename std::conditional<WT, fvec<KPL>, ivec<KPL>>::type;
    __shared__ double qa[128], qb[128];
    const int lane = lane_id();
    const C* nkw = reinterpret_cast<const C*>(p.nkw);
    const C* cells = reinterpret_cast<const C*>(p.cells);
    const int64_t cb = p.counter ? p.chunk_base[0] : p.chunk_base[p.job_base];
    const int64_t ce = p.counter ? p.chunk_base[p.njobs] : p.chunk_base[p.job_base + 1];
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int64_t ch = cb + (int64_t)blockIdx.x; ch < ce; ch += (int64_t)gridDim.x) {
        const int64_t rb = p.chunk_rb[ch], re = p.chunk_rb[ch + 1];
        double acc = 0.0;
        int qn = 0;
        CV m0, m1, n0, n1;  // this run's model and private columns, and the next run's
        auto fetch = [&](int64_t r, CV& x0, CV& x1) {
            const uint32_t v = p.run_word[r];
            if constexpr (WT) {
                load_kvec_f<KPL>(nkw + (int64_t)v * p.Kp, lane, x0);
                load_kvec_f<KPL>(cells + r * p.Kp, lane, x1);
            } else {
                load_kvec<KPL>(nkw + (int64_t)v * p.Kp, lane, x0);
                load_kvec<KPL>(cells + r * p.Kp, lane, x1);
            }
        };
        if (rb < re) fetch(rb, n0, n1);
        for (int64_t r = rb; r < re; ++r) {
            m0 = n0;
            m1 = n1;
            if (r + 1 < re) fetch(r + 1, n0, n1);
            int pr = -1;
            if constexpr (ASYM) pr = p.prior_row[p.run_word[r]];
#pragma unroll
            for (int j = 0; j < KPL; ++j) {
                const C c0 = m0[j], c1 = m1[j];
                const bool d = lane * KPL + j < p.K && c0 != c1;
                const unsigned long long m = __ballot(d);
                if (!m) continue;
                if (d) {
                    const int q = kpos<KPL>(lane, j);
                    const double e = ASYM && pr >= 0 ? (double)p.prior_eta[(int64_t)pr * p.Kp + q] : p.eta_d;
                    const int at = qn + __popcll(m & below);
                    qa[at] = c1 != (C)0 ? (double)c1 + e : e;
                    qb[at] = c0 != (C)0 ? (double)c0 + e : e;
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
        __syncthreads();
        if (lane < qn) acc += lgamma(qa[lane]) - lgamma(qb[lane]);
        acc = wave_sum_d(acc);
        if (lane == 0) p.chunk_ll[ch] = acc;
        __syncthreads();  // the queue is read before the next chunk writes it
    }
}

template <int KPL, bool WT, bool ASYM>
__global__ void __launch_bounds__(64) infer_gen_docs_kernel(InferGenParams p) {
    using C = typename std::conditional<WT, float, int32_t>::type;
    const int lane = lane_id();
    for (int64_t idx = p.counter ? (int64_t)blockIdx.x : p.job_base; idx < p.njobs;
         idx += p.counter ? (int64_t)gridDim.x : p.njobs) {
        const InferJob J = p.jobs[idx];
        const C* nd = reinterpret_cast<const C*>(p.ndk_out) + (int64_t)J.slot * p.Kp;
        const C* nk1 = reinterpret_cast<const C*>(p.nk_out) + (int64_t)J.slot * p.Kp;
        const C* nk0 = reinterpret_cast<const C*>(p.nk);
        const float den = p.docw[J.slot] + p.asum_f;
        double ll = 0.0;
#pragma unroll
        for (int j = 0; j < KPL; ++j) {
            const int k = lane * KPL + j;
            if (k >= p.K) continue;
            const int q = kpos<KPL>(lane, j);
            const float a = p.alpha[q];
            const C x = nd[q];
            p.theta[(int64_t)J.slot * p.K + k] = ((float)x + a) / den;
            if (x != (C)0) ll += lgamma((double)x + (double)a) - lgamma((double)a);
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

}  // namespace

// ================================================================ launchers
#define LDA_DISPATCH_KPL(kpl, CALL)            \
    switch (kpl) {                             \
        case 1: { constexpr int KPL = 1; CALL; } break;   \
        case 2: { constexpr int KPL = 2; CALL; } break;   \
        case 4: { constexpr int KPL = 4; CALL; } break;   \
        case 8: { constexpr int KPL = 8; CALL; } break;   \
        case 16: { constexpr int KPL = 16; CALL; } break; \
        case 32: { constexpr int KPL = 32; CALL; } break; \
        default: return hipErrorInvalidValue;             \
    }

// ASYM: the word-prior variant of the training kernels (SweepParams.prior_row set)
#define LDA_DISPATCH_KPL_ETA(kpl, asym, CALL)                          \
    if (asym) {                                                        \
        constexpr bool ASYM = true;                                    \
        switch (kpl) {                                                 \
            case 1: { constexpr int KPL = 1; CALL; } break;            \
            case 2: { constexpr int KPL = 2; CALL; } break;            \
            case 4: { constexpr int KPL = 4; CALL; } break;            \
            case 8: { constexpr int KPL = 8; CALL; } break;            \
            case 16: { constexpr int KPL = 16; CALL; } break;          \
            case 32: { constexpr int KPL = 32; CALL; } break;          \
            default: return hipErrorInvalidValue;                      \
        }                                                              \
    } else {                                                           \
        constexpr bool ASYM = false;                                   \
        switch (kpl) {                                                 \
            case 1: { constexpr int KPL = 1; CALL; } break;            \
            case 2: { constexpr int KPL = 2; CALL; } break;            \
            case 4: { constexpr int KPL = 4; CALL; } break;            \
            case 8: { constexpr int KPL = 8; CALL; } break;            \
            case 16: { constexpr int KPL = 16; CALL; } break;          \
            case 32: { constexpr int KPL = 32; CALL; } break;          \
            default: return hipErrorInvalidValue;                      \
        }                                                              \
    }

hipError_t launch_init_topics(const uint32_t* words, uint16_t* z, uint64_t N, const float* ca, int K, int V, uint32_t s0,
                              uint32_t s1, const int32_t* prior_row, const float* prior_eta, int kpl, int Kp, float eta,
                              hipStream_t s) {
    if (!N) return hipSuccess;
    uint64_t blocks = (N + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(init_topics_kernel, dim3((unsigned)blocks), dim3(256), 0, s, words, z, N, ca, K, V, s0, s1,
                       prior_row, prior_eta, kpl, Kp, eta);
    return hipGetLastError();
}

hipError_t launch_build_counts(int kpl, const uint32_t* words, const uint16_t* z, const uint64_t* off, int64_t D,
                               int32_t* ndk, int32_t* nkw, int32_t* nk, int Kp, int V, hipStream_t s) {
    if (!D) return hipSuccess;
    int64_t blocks = (D + 3) / 4;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(build_counts_kernel, dim3((unsigned)blocks), dim3(256), 0, s, words, z, off, D, ndk, nkw, nk, Kp,
                       V, kpl);
    return hipGetLastError();
}

hipError_t launch_sweep_device(int kpl, const SweepParams& p, int grid, int waves, hipStream_t s) {
    if (waves < 1 || waves > kSweepWaves) return hipErrorInvalidValue;
    if (kpl > kSmallMaxKpl || p.exact)
        return !p.nkw16 ? launch_bk_sweep_device(kpl, p, false, grid, waves, s) : hipErrorInvalidValue;
    if (p.nkw16) {
        LDA_DISPATCH_KPL_ETA(kpl, p.prior_row != nullptr,
                             hipLaunchKernelGGL((sweep_device_kernel<KPL, ASYM, true>), dim3(grid), dim3(64 * waves),
                                                wave_lds_bytes<KPL>(waves), s, p));
    } else {
        LDA_DISPATCH_KPL_ETA(kpl, p.prior_row != nullptr,
                             hipLaunchKernelGGL((sweep_device_kernel<KPL, ASYM, false>), dim3(grid), dim3(64 * waves),
                                                wave_lds_bytes<KPL>(waves), s, p));
    }
    return hipGetLastError();
}

hipError_t launch_sweep_copy(int kpl, const SweepParams& p, hipStream_t s) {
    if (kpl > kSmallMaxKpl || p.exact) return launch_bk_sweep_copy(kpl, p, false, s);
    LDA_DISPATCH_KPL_ETA(kpl, p.prior_row != nullptr,
                         hipLaunchKernelGGL((sweep_copy_kernel<KPL, ASYM>), dim3(p.W), dim3(64), 0, s, p));
    return hipGetLastError();
}

hipError_t launch_merge_copies(int32_t* nkw, const int32_t* loc, int64_t n, int W, hipStream_t s) {
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(merge_copies_kernel, dim3((unsigned)blocks), dim3(256), 0, s, nkw, loc, n, W);
    return hipGetLastError();
}

hipError_t launch_replicate(int32_t* dst, const int32_t* src, int64_t n, int W, hipStream_t s) {
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(replicate_kernel, dim3((unsigned)blocks), dim3(256), 0, s, dst, src, n, W);
    return hipGetLastError();
}

hipError_t launch_sweep_det(int kpl, const SweepParams& p, hipStream_t s) {
    if (kpl > kSmallMaxKpl || p.exact) return launch_bk_sweep_det(kpl, p, false, s);
    LDA_DISPATCH_KPL_ETA(kpl, p.prior_row != nullptr,
                         hipLaunchKernelGGL((sweep_det_kernel<KPL, ASYM>), dim3(p.W), dim3(64), 0, s, p));
    return hipGetLastError();
}

template <int KPL, bool ASYM>
static int sweep_device_blocks_kpl(int waves) {
    int n = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sweep_device_kernel<KPL, ASYM, false>,
                                                                      64 * waves, wave_lds_bytes<KPL>(waves));
    return (e == hipSuccess && n > 0) ? n : 1;
}

int sweep_device_blocks_per_cu(int kpl, bool asym, int waves) {
    if (kpl > kSmallMaxKpl) return waves == 1 ? std::max(bk_sweep_blocks_per_cu(false, asym, true, kpl), 1) : 0;
    switch (kpl) {
        case 1: return asym ? sweep_device_blocks_kpl<1, true>(waves) : sweep_device_blocks_kpl<1, false>(waves);
        case 2: return asym ? sweep_device_blocks_kpl<2, true>(waves) : sweep_device_blocks_kpl<2, false>(waves);
        case 4: return asym ? sweep_device_blocks_kpl<4, true>(waves) : sweep_device_blocks_kpl<4, false>(waves);
        case 8: return asym ? sweep_device_blocks_kpl<8, true>(waves) : sweep_device_blocks_kpl<8, false>(waves);
        case 16: return asym ? sweep_device_blocks_kpl<16, true>(waves) : sweep_device_blocks_kpl<16, false>(waves);
        case 32: return asym ? sweep_device_blocks_kpl<32, true>(waves) : sweep_device_blocks_kpl<32, false>(waves);
        default: return 1;
    }
}

template <int KPL, bool ASYM>
static int infer_blocks_kpl(int segs) {
    constexpr size_t lds = (64 * KPL + 12) * sizeof(int);
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (segs) {
        case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, infer_kernel<KPL, 1, ASYM>, 64, 0); break;
        case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, infer_kernel<KPL, 2, ASYM>, 128, lds); break;
        case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, infer_kernel<KPL, 4, ASYM>, 256, lds); break;
        case 8: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, infer_kernel<KPL, 8, ASYM>, 512, lds); break;
        default: break;
    }
    return e == hipSuccess ? n : 0;
}

int infer_blocks_per_cu(int kpl, int segs, bool asym) {
    switch (kpl) {
        case 1: return asym ? infer_blocks_kpl<1, true>(segs) : infer_blocks_kpl<1, false>(segs);
        case 2: return asym ? infer_blocks_kpl<2, true>(segs) : infer_blocks_kpl<2, false>(segs);
        case 4: return asym ? infer_blocks_kpl<4, true>(segs) : infer_blocks_kpl<4, false>(segs);
        case 8: return asym ? infer_blocks_kpl<8, true>(segs) : infer_blocks_kpl<8, false>(segs);
        case 16: return asym ? infer_blocks_kpl<16, true>(segs) : infer_blocks_kpl<16, false>(segs);
        case 32: return asym ? infer_blocks_kpl<32, true>(segs) : infer_blocks_kpl<32, false>(segs);
        default: return 0;
    }
}

hipError_t launch_topic_word_dist(const void* nkw, bool fl, int V, int Kp, int pos, float eta, float sum, int normalize,
                                  float* out, hipStream_t s) {
    const int blocks = std::max(1, std::min((V + 255) / 256, 4096));
    if (fl)
        hipLaunchKernelGGL(topic_word_dist_kernel<float>, dim3(blocks), dim3(256), 0, s,
                           reinterpret_cast<const float*>(nkw), V, Kp, pos, eta, sum, normalize, out);
    else
        hipLaunchKernelGGL(topic_word_dist_kernel<int32_t>, dim3(blocks), dim3(256), 0, s,
                           reinterpret_cast<const int32_t*>(nkw), V, Kp, pos, eta, sum, normalize, out);
    return hipGetLastError();
}

hipError_t launch_count_topics(const uint32_t* words, const uint16_t* z, uint64_t N, int V, int K,
                               unsigned long long* cnt, hipStream_t s) {
    const int blocks = (int)std::max<uint64_t>(1, std::min<uint64_t>((N + 255) / 256, 1024));
    hipLaunchKernelGGL(count_topics_kernel, dim3(blocks), dim3(256), K * sizeof(int), s, words, z, N, V, K, cnt);
    return hipGetLastError();
}

hipError_t launch_probe_read(const void* src, int64_t bytes, int* sink, int grid, hipStream_t s) {
    hipLaunchKernelGGL(probe_read_kernel, dim3(grid), dim3(256), 0, s, reinterpret_cast<const int4*>(src), bytes / 16, sink);
    return hipGetLastError();
}

hipError_t launch_probe_copy(const void* src, void* dst, int64_t bytes, int grid, hipStream_t s) {
    hipLaunchKernelGGL(probe_copy_kernel, dim3(grid), dim3(256), 0, s, reinterpret_cast<const int4*>(src),
                       reinterpret_cast<int4*>(dst), bytes / 16);
    return hipGetLastError();
}

hipError_t launch_pack_narrow(const int32_t* nkw, uint16_t* nkw16, int hot, int V, int Kp, uint32_t* bad, hipStream_t s) {
    const int64_t n = (int64_t)(V - hot) * Kp;
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(pack_narrow_kernel, dim3((unsigned)blocks), dim3(256), 0, s, nkw, nkw16, (int64_t)hot * Kp, n, bad);
    return hipGetLastError();
}

hipError_t launch_unpack_narrow(int32_t* nkw, const uint16_t* nkw16, int hot, int V, int Kp, hipStream_t s) {
    const int64_t n = (int64_t)(V - hot) * Kp;
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(unpack_narrow_kernel, dim3((unsigned)blocks), dim3(256), 0, s, nkw, nkw16, (int64_t)hot * Kp, n);
    return hipGetLastError();
}

hipError_t launch_copy_i32(int32_t* dst, const int32_t* src, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(copy_i32_kernel, dim3((unsigned)blocks), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_add_i32(int32_t* dst, const int32_t* src, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(add_i32_kernel, dim3((unsigned)blocks), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* dst, uint32_t v, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)blocks), dim3(256), 0, s, dst, v, n);
    return hipGetLastError();
}

hipError_t launch_rowsum_nkw(const int32_t* nkw, int V, int Kp, int32_t* nk, hipStream_t s) {
    hipError_t e = launch_fill_u32(reinterpret_cast<uint32_t*>(nk), 0u, Kp, s);
    if (e != hipSuccess || V <= 0) return e;
    const int words = std::max(kRowsumWords, (V + 65534) / 65535);
    const unsigned chunks = (unsigned)((V + words - 1) / words);
    hipLaunchKernelGGL(rowsum_chunks_kernel, dim3((Kp + 255) / 256, chunks), dim3(256), 0, s, nkw, V, Kp, words, nk);
    return hipGetLastError();
}

hipError_t launch_loglik(const int32_t* ndk, const int32_t* doclen, int64_t D, const int32_t* nkw, const int32_t* nk,
                         const float* alpha, int V, int Kp, double asum, double eta, double veta, double* partial,
                         int nblk, const int32_t* prior_row, const float* prior_eta, const double* etasum_d,
                         hipStream_t s) {
    hipLaunchKernelGGL(loglik_kernel<int32_t>, dim3(nblk), dim3(256), 0, s, ndk, doclen, D, nkw, nk, alpha, V, Kp, asum,
                       eta, veta, partial, prior_row, prior_eta, etasum_d);
    return hipGetLastError();
}

hipError_t launch_loglik_f(const float* ndk, const float* docw, int64_t D, const float* nkw, const float* nk,
                           const float* alpha, int V, int Kp, double asum, double eta, double veta, double* partial,
                           int nblk, const int32_t* prior_row, const float* prior_eta, const double* etasum_d,
                           hipStream_t s) {
    hipLaunchKernelGGL(loglik_kernel<float>, dim3(nblk), dim3(256), 0, s, ndk, docw, D, nkw, nk, alpha, V, Kp, asum,
                       eta, veta, partial, prior_row, prior_eta, etasum_d);
    return hipGetLastError();
}

hipError_t launch_alpha_sums(const int32_t* ndk, const int32_t* doclen, int64_t D, const float* alpha, int Kp,
                             const float* asum, double* part, int nblk, hipStream_t s) {
    hipLaunchKernelGGL(alpha_sums_kernel<int32_t>, dim3(nblk), dim3(256), 0, s, ndk, doclen, D, alpha, Kp, asum, part);
    return hipGetLastError();
}

hipError_t launch_alpha_sums_f(const float* ndk, const float* docw, int64_t D, const float* alpha, int Kp,
                               const float* asum, double* part, int nblk, hipStream_t s) {
    hipLaunchKernelGGL(alpha_sums_kernel<float>, dim3(nblk), dim3(256), 0, s, ndk, docw, D, alpha, Kp, asum, part);
    return hipGetLastError();
}

hipError_t launch_alpha_asum(const float* alpha, int K, int kpl, float* out, hipStream_t s) {
    hipLaunchKernelGGL(alpha_asum_kernel, dim3(1), dim3(64), 0, s, alpha, K, kpl, out);
    return hipGetLastError();
}

hipError_t launch_alpha_update_direct(float* alpha, const double* part, int nblk, int Kp, int64_t D, const float* asum,
                                      const float* dglut, float amax, unsigned* bad, hipStream_t s) {
    hipLaunchKernelGGL(alpha_update_direct_kernel, dim3(Kp), dim3(256), 0, s, alpha, part, nblk, Kp, D, asum, dglut, amax,
                       bad);
    return hipGetLastError();
}

hipError_t launch_alpha_hist(const int32_t* ndk, const int32_t* doclen, int64_t D, int Kp, int Lmax, int32_t* hist,
                             int32_t* hist_len, hipStream_t s) {
    int64_t blocks = (D * (int64_t)Kp + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(alpha_hist_kernel, dim3((unsigned)blocks), dim3(256), 0, s, ndk, doclen, D, Kp, Lmax, hist, hist_len);
    return hipGetLastError();
}

hipError_t launch_alpha_den_hist(const int32_t* hist_len, int Lmax, int64_t D, const float* asum, const float* dglut,
                                 float* den, hipStream_t s) {
    hipLaunchKernelGGL(alpha_den_hist_kernel, dim3(1), dim3(256), 0, s, hist_len, Lmax, D, asum, dglut, den);
    return hipGetLastError();
}

hipError_t launch_alpha_update_hist(float* alpha, const int32_t* hist, int Lmax, int Kp, int64_t D, const float* den,
                                    const float* dglut, float amax, unsigned* bad, hipStream_t s) {
    hipLaunchKernelGGL(alpha_update_hist_kernel, dim3(Kp), dim3(256), 0, s, alpha, hist, Lmax, D, den, dglut, amax, bad);
    return hipGetLastError();
}

hipError_t launch_sweep_device_w(int kpl, const SweepParams& p, int grid, int waves, hipStream_t s) {
    if (waves < 1 || waves > kSweepWaves) return hipErrorInvalidValue;
    if (kpl > kSmallMaxKpl) return launch_bk_sweep_device(kpl, p, true, grid, waves, s);
    LDA_DISPATCH_KPL_ETA(kpl, p.prior_row != nullptr,
                         hipLaunchKernelGGL((sweep_device_w_kernel<KPL, ASYM>), dim3(grid), dim3(64 * waves),
                                            wave_lds_bytes<KPL>(waves), s, p));
    return hipGetLastError();
}

hipError_t launch_sweep_det_w(int kpl, const SweepParams& p, hipStream_t s) {
    if (kpl > kSmallMaxKpl) return launch_bk_sweep_det(kpl, p, true, s);
    LDA_DISPATCH_KPL_ETA(kpl, p.prior_row != nullptr,
                         hipLaunchKernelGGL((sweep_det_w_kernel<KPL, ASYM>), dim3(p.W), dim3(64), 0, s, p));
    return hipGetLastError();
}

hipError_t launch_sweep_copy_w(int kpl, const SweepParams& p, hipStream_t s) {
    if (kpl > kSmallMaxKpl) return launch_bk_sweep_copy(kpl, p, true, s);
    LDA_DISPATCH_KPL_ETA(kpl, p.prior_row != nullptr,
                         hipLaunchKernelGGL((sweep_copy_w_kernel<KPL, ASYM>), dim3(p.W), dim3(64), 0, s, p));
    return hipGetLastError();
}

hipError_t launch_merge_copies_f(float* nkw, const float* loc, int64_t n, int W, hipStream_t s) {
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(merge_copies_f_kernel, dim3((unsigned)blocks), dim3(256), 0, s, nkw, loc, n, W);
    return hipGetLastError();
}

template <int KPL, bool ASYM>
static int sweep_det_blocks_kpl(bool weighted) {
    int n = 0;
    const hipError_t e =
        weighted ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sweep_det_w_kernel<KPL, ASYM>, 64, 0)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sweep_det_kernel<KPL, ASYM>, 64, 0);
    return (e == hipSuccess && n > 0) ? n : 1;
}

int sweep_det_blocks_per_cu(int kpl, bool asym, bool weighted) {
    if (kpl > kSmallMaxKpl) return std::max(bk_sweep_blocks_per_cu(weighted, asym, false, kpl), 1);
    switch (kpl) {
        case 1: return asym ? sweep_det_blocks_kpl<1, true>(weighted) : sweep_det_blocks_kpl<1, false>(weighted);
        case 2: return asym ? sweep_det_blocks_kpl<2, true>(weighted) : sweep_det_blocks_kpl<2, false>(weighted);
        case 4: return asym ? sweep_det_blocks_kpl<4, true>(weighted) : sweep_det_blocks_kpl<4, false>(weighted);
        case 8: return asym ? sweep_det_blocks_kpl<8, true>(weighted) : sweep_det_blocks_kpl<8, false>(weighted);
        case 16: return asym ? sweep_det_blocks_kpl<16, true>(weighted) : sweep_det_blocks_kpl<16, false>(weighted);
        case 32: return asym ? sweep_det_blocks_kpl<32, true>(weighted) : sweep_det_blocks_kpl<32, false>(weighted);
        default: return 1;
    }
}

template <int KPL, bool ASYM>
static int sweep_device_w_blocks_kpl(int waves) {
    int n = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, sweep_device_w_kernel<KPL, ASYM>,
                                                                      64 * waves, wave_lds_bytes<KPL>(waves));
    return (e == hipSuccess && n > 0) ? n : 1;
}

int sweep_device_w_blocks_per_cu(int kpl, bool asym, int waves) {
    if (kpl > kSmallMaxKpl) return waves == 1 ? std::max(bk_sweep_blocks_per_cu(true, asym, true, kpl), 1) : 0;
    switch (kpl) {
        case 1: return asym ? sweep_device_w_blocks_kpl<1, true>(waves) : sweep_device_w_blocks_kpl<1, false>(waves);
        case 2: return asym ? sweep_device_w_blocks_kpl<2, true>(waves) : sweep_device_w_blocks_kpl<2, false>(waves);
        case 4: return asym ? sweep_device_w_blocks_kpl<4, true>(waves) : sweep_device_w_blocks_kpl<4, false>(waves);
        case 8: return asym ? sweep_device_w_blocks_kpl<8, true>(waves) : sweep_device_w_blocks_kpl<8, false>(waves);
        case 16: return asym ? sweep_device_w_blocks_kpl<16, true>(waves) : sweep_device_w_blocks_kpl<16, false>(waves);
        case 32: return asym ? sweep_device_w_blocks_kpl<32, true>(waves) : sweep_device_w_blocks_kpl<32, false>(waves);
        default: return 1;
    }
}

__global__ void selftest_reciprocal_kernel(int exp_shift, unsigned long long* bad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (1u << 23)) return;
    const float b = __uint_as_float(((uint32_t)(127 + exp_shift) << 23) | i);
    if (__float_as_uint(recip_rn(b)) != __float_as_uint(1.0f / b)) atomicAdd(bad, 1ull);
}

hipError_t launch_selftest_reciprocal(int exp_shift, unsigned long long* bad, hipStream_t s) {
    hipLaunchKernelGGL(selftest_reciprocal_kernel, dim3((1u << 23) / 256), dim3(256), 0, s, exp_shift, bad);
    return hipGetLastError();
}

hipError_t launch_clamp_rowsum_f(float* nkw, int V, int Kp, float* part, float* nk, hipStream_t s) {
    const int nb = (V + kRowsumBlock - 1) / kRowsumBlock;
    if (nb > 0) {
        hipLaunchKernelGGL(clamp_rowsum_f_kernel, dim3(nb, (Kp + 255) / 256), dim3(256), 0, s, nkw, V, Kp, part);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(rowsum_blocks_f_kernel, dim3((Kp + 255) / 256), dim3(256), 0, s, part, nb, Kp, nk);
    return hipGetLastError();
}

hipError_t launch_doclen(const uint32_t* words, const uint64_t* off, int64_t D, int V, int32_t* doclen, hipStream_t s) {
    if (!D) return hipSuccess;
    int64_t blocks = (D + 3) / 4;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(doclen_kernel, dim3((unsigned)blocks), dim3(256), 0, s, words, off, D, V, doclen);
    return hipGetLastError();
}

hipError_t launch_invariants(int kpl, const uint32_t* words, const uint16_t* z, const uint64_t* off, int64_t D,
                             const int32_t* ndk, const int32_t* nkw, const int32_t* nk, const int32_t* doclen, int K,
                             int V, int Kp, int32_t* sndk, int32_t* snkw, long long* colsum, unsigned long long* flags,
                             hipStream_t s) {
    (void)nk;
    int64_t blocks = (D + 3) / 4;
    if (blocks > 16384) blocks = 16384;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(recount_kernel, dim3((unsigned)blocks), dim3(256), 0, s, words, z, off, D, K, V, Kp, sndk, snkw,
                       flags, kpl);
    hipLaunchKernelGGL(compare_kernel, dim3(1024), dim3(256), 0, s, ndk, sndk, D * (int64_t)Kp, flags + 3, flags + 5);
    hipLaunchKernelGGL(compare_kernel, dim3(1024), dim3(256), 0, s, nkw, snkw, (int64_t)V * Kp, flags + 4, flags + 5);
    hipLaunchKernelGGL(ndk_rows_kernel, dim3((unsigned)blocks), dim3(256), 0, s, ndk, doclen, D, Kp, flags + 0);
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((2 * (int64_t)Kp + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<uint32_t*>(colsum), 0u, 2 * (int64_t)Kp);
    if (V > 0) {
        const int words = std::max(kRowsumWords, (V + 65534) / 65535);
        const unsigned chunks = (unsigned)((V + words - 1) / words);
        hipLaunchKernelGGL(colsum_chunks_kernel, dim3((Kp + 255) / 256, chunks), dim3(256), 0, s, nkw, V, Kp, words,
                           colsum);
    }
    return hipGetLastError();
}

hipError_t launch_sample_frozen(int kpl, const int32_t* ndk, const int32_t* nkw, const int32_t* nk, const float* alpha,
                                float eta, float veta, int K, int V, int Kp, int64_t n, const int64_t* doc,
                                const uint32_t* word, const uint16_t* z_old, const float* u, uint16_t* z_new,
                                float* total, int32_t* bk_scratch, int bk_grid, hipStream_t s) {
    (void)V;
    if (n <= 0) return hipSuccess;
    if (kpl > kSmallMaxKpl || bk_scratch)
        return launch_bk_sample_frozen(kpl, ndk, nkw, nk, alpha, eta, veta, K, Kp, n, bk_grid, doc, word, z_old, u, z_new,
                                       total, bk_scratch, s);
    int64_t grid = n < 65536 ? n : 65536;
    LDA_DISPATCH_KPL(kpl, hipLaunchKernelGGL(sample_frozen_kernel<KPL>, dim3((unsigned)grid), dim3(64), wave_lds_bytes<KPL>(), s, ndk, nkw, nk,
                                             alpha, eta, veta, K, Kp, n, doc, word, z_old, u, z_new, total));
    return hipGetLastError();
}

template <int KPL, bool ASYM>
static hipError_t launch_infer_kpl(int segs, const InferParams* p_dev, int grid, hipStream_t s) {
    constexpr size_t lds = (64 * KPL + 12) * sizeof(int);
    switch (segs) {
        case 1: hipLaunchKernelGGL((infer_kernel<KPL, 1, ASYM>), dim3(grid), dim3(64), 0, s, p_dev); break;
        case 2: hipLaunchKernelGGL((infer_kernel<KPL, 2, ASYM>), dim3(grid), dim3(128), lds, s, p_dev); break;
        case 4: hipLaunchKernelGGL((infer_kernel<KPL, 4, ASYM>), dim3(grid), dim3(256), lds, s, p_dev); break;
        case 8: hipLaunchKernelGGL((infer_kernel<KPL, 8, ASYM>), dim3(grid), dim3(512), lds, s, p_dev); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_infer(int kpl, int segs, bool asym, const InferParams* p_dev, int grid, hipStream_t s) {
    if (asym) {
        LDA_DISPATCH_KPL(kpl, return (launch_infer_kpl<KPL, true>(segs, p_dev, grid, s)));
    } else {
        LDA_DISPATCH_KPL(kpl, return (launch_infer_kpl<KPL, false>(segs, p_dev, grid, s)));
    }
    return hipErrorInvalidValue;
}

template <int KPL, bool WT, bool ASYM>
static hipError_t launch_infer_gen_k(int segs, const InferGenParams& p, int grid, bool finalize, hipStream_t s) {
    if (finalize) {
        hipLaunchKernelGGL((infer_gen_runs_kernel<KPL, WT, ASYM>), dim3(grid), dim3(64), 0, s, p);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((infer_gen_docs_kernel<KPL, WT, ASYM>), dim3(grid), dim3(64), 0, s, p);
    } else if (WT && !ASYM && KPL <= kFoldWMaxKpl && segs == 1)
        hipLaunchKernelGGL((infer_w_kernel<KPL, false>), dim3(grid), dim3(64), 0, s, p);
    else if (WT && !ASYM && KPL <= kFoldWMaxKpl)
        hipLaunchKernelGGL((infer_w_kernel<KPL, true>), dim3(grid), dim3(64 * segs), gen_w_lds_bytes<KPL>(segs), s, p);
    else if (segs == 1)
        hipLaunchKernelGGL((infer_gen_kernel<KPL, WT, ASYM, false>), dim3(grid), dim3(64), gen_lds_bytes<KPL>(1), s, p);
    else
        hipLaunchKernelGGL((infer_gen_kernel<KPL, WT, ASYM, true>), dim3(grid), dim3(64 * segs), gen_lds_bytes<KPL>(segs),
                           s, p);
    return hipGetLastError();
}

template <bool WT, bool ASYM>
static hipError_t launch_infer_gen_t(int kpl, int segs, const InferGenParams& p, int grid, bool finalize, hipStream_t s) {
    switch (kpl) {
        case 1: return launch_infer_gen_k<1, WT, ASYM>(segs, p, grid, finalize, s);
        case 2: return launch_infer_gen_k<2, WT, ASYM>(segs, p, grid, finalize, s);
        case 4: return launch_infer_gen_k<4, WT, ASYM>(segs, p, grid, finalize, s);
        case 8: return launch_infer_gen_k<8, WT, ASYM>(segs, p, grid, finalize, s);
        case 16: return launch_infer_gen_k<16, WT, ASYM>(segs, p, grid, finalize, s);
        case 32: return launch_infer_gen_k<32, WT, ASYM>(segs, p, grid, finalize, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_infer_gen(int kpl, int segs, const InferGenParams& p, int grid, bool finalize, hipStream_t s) {
    if (segs != 1 && segs != 2 && segs != 4 && segs != kMaxFoldSegments) return hipErrorInvalidValue;
    if (kpl > kSmallMaxKpl || p.exact)
        return segs == 1 ? launch_bk_infer_gen(kpl, p, grid, finalize, s) : hipErrorInvalidValue;
    const bool wt = p.wt != nullptr, asym = p.prior_row != nullptr;
    if (wt) return asym ? launch_infer_gen_t<true, true>(kpl, segs, p, grid, finalize, s)
                        : launch_infer_gen_t<true, false>(kpl, segs, p, grid, finalize, s);
    return asym ? launch_infer_gen_t<false, true>(kpl, segs, p, grid, finalize, s)
                : launch_infer_gen_t<false, false>(kpl, segs, p, grid, finalize, s);
}

template <int KPL, bool SEG>
static int infer_gen_blocks_seg(bool wt, bool asym, int segs, size_t lds) {
    int n = 0;
    hipError_t e;
    if (wt) e = asym ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, infer_gen_kernel<KPL, true, true, SEG>, 64 * segs, lds)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, infer_gen_kernel<KPL, true, false, SEG>, 64 * segs, lds);
    else e = asym ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, infer_gen_kernel<KPL, false, true, SEG>, 64 * segs, lds)
                  : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, infer_gen_kernel<KPL, false, false, SEG>, 64 * segs, lds);
    return e == hipSuccess ? n : 0;
}

template <int KPL>
static int infer_gen_blocks_kpl(bool wt, bool asym, int segs) {
    if (wt && !asym && KPL <= kFoldWMaxKpl) {
        // the register-resident kernel (infer_w_kernel)
        const size_t lds = gen_w_lds_bytes<KPL>(segs);
        if (lds > 160 * 1024) return 0;
        int n = 0;
        const hipError_t e = segs == 1
            ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, infer_w_kernel<KPL, false>, 64, lds)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, infer_w_kernel<KPL, true>, 64 * segs, lds);
        return e == hipSuccess ? n : 0;
    }
    const size_t lds = gen_lds_bytes<KPL>(segs);
    if (lds > 160 * 1024) return 0;
    return segs == 1 ? infer_gen_blocks_seg<KPL, false>(wt, asym, 1, lds) : infer_gen_blocks_seg<KPL, true>(wt, asym, segs, lds);
}

int infer_gen_blocks_per_cu(int kpl, bool wt, bool asym, int segs) {
    if (kpl > kSmallMaxKpl) return segs == 1 ? bk_infer_gen_blocks_per_cu(wt, asym, kpl) : 0;
    switch (kpl) {
        case 1: return infer_gen_blocks_kpl<1>(wt, asym, segs);
        case 2: return infer_gen_blocks_kpl<2>(wt, asym, segs);
        case 4: return infer_gen_blocks_kpl<4>(wt, asym, segs);
        case 8: return infer_gen_blocks_kpl<8>(wt, asym, segs);
        case 16: return infer_gen_blocks_kpl<16>(wt, asym, segs);
        case 32: return infer_gen_blocks_kpl<32>(wt, asym, segs);
        default: return 0;
    }
}

template <int KPL, bool ASYM>
static hipError_t launch_infer_finalize_k(const InferParams* p_dev, int grid, hipStream_t s) {
    hipLaunchKernelGGL((infer_finalize_kernel<KPL, ASYM>), dim3(grid), dim3(64), wave_lds_bytes<KPL>(), s, p_dev);
    return hipGetLastError();
}

hipError_t launch_infer_finalize(int kpl, bool asym, const InferParams* p_dev, int grid, hipStream_t s) {
    if (asym) {
        LDA_DISPATCH_KPL(kpl, return (launch_infer_finalize_k<KPL, true>(p_dev, grid, s)));
    } else {
        LDA_DISPATCH_KPL(kpl, return (launch_infer_finalize_k<KPL, false>(p_dev, grid, s)));
    }
    return hipErrorInvalidValue;
}

}  // namespace lda
