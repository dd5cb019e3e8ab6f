// lda_kernels_common.h -- device primitives shared by the gfx950 kernel
// translation units (lda_device.hip, lda_bigk.hip): the Philox stream, the
// reciprocal, lane-layout loads and stores, the DPP wave scan, the lane-local
// inverse-CDF count, forShuffled's prime table and small FP helpers.  Every
// definition has internal linkage, so each unit carries its own copy.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lda_internal.h"

namespace lda {
namespace {

constexpr uint32_t PHILOX_M0 = 0xD2511F53u, PHILOX_M1 = 0xCD9E8D57u;
constexpr uint32_t PHILOX_W0 = 0x9E3779B9u, PHILOX_W1 = 0xBB67AE85u;

// Philox4x32-10 (Salmon et al., SC'11), first output word.
__device__ __forceinline__ uint32_t philox_x0(uint32_t k0, uint32_t k1, uint32_t c0, uint32_t c1, uint32_t c2,
                                              uint32_t c3) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = __umulhi(PHILOX_M0, c0), lo0 = PHILOX_M0 * c0;
        uint32_t hi1 = __umulhi(PHILOX_M1, c2), lo1 = PHILOX_M1 * c2;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += PHILOX_W0; k1 += PHILOX_W1;
    }
    return c0;
}

__device__ __forceinline__ float uniform24(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

// RN(1 / b) for normal b: the hardware reciprocal refined by one Newton step,
// e = 1 - b*r0, r = r0 + e*r0 (both fused).  The refinement returns the
// correctly rounded reciprocal for every float mantissa on gfx950 (all 2^23
// mantissas, checked by lda_selftest_reciprocal), and the hardware
// approximation depends on the mantissa only, so the result equals the IEEE
// division 1.0f / b at a fraction of its instructions.
__device__ __forceinline__ float recip_rn(float b) {
    const float r0 = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, r0, 1.0f);
    return __builtin_fmaf(e, r0, r0);
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63u); }

template <int KPL>
__device__ __forceinline__ int kpos(int lane, int j) {
    if constexpr (KPL < 4) return lane * KPL + j;
    else return (j >> 2) * 256 + lane * 4 + (j & 3);
}

template <int KPL, typename V>
__device__ __forceinline__ void load_kvec(const int32_t* __restrict__ base, int lane, V& o) {
    if constexpr (KPL >= 4) {
#pragma unroll
        for (int g = 0; g < KPL / 4; ++g) {
            int4 x = *reinterpret_cast<const int4*>(base + g * 256 + lane * 4);
            o[4 * g] = x.x; o[4 * g + 1] = x.y; o[4 * g + 2] = x.z; o[4 * g + 3] = x.w;
        }
    } else if constexpr (KPL == 2) {
        int2 x = *reinterpret_cast<const int2*>(base + lane * 2);
        o[0] = x.x; o[1] = x.y;
    } else {
        o[0] = base[lane];
    }
}

template <int KPL, typename V>
__device__ __forceinline__ void load_kvec_f(const float* __restrict__ base, int lane, V& o) {
    if constexpr (KPL >= 4) {
#pragma unroll
        for (int g = 0; g < KPL / 4; ++g) {
            float4 x = *reinterpret_cast<const float4*>(base + g * 256 + lane * 4);
            o[4 * g] = x.x; o[4 * g + 1] = x.y; o[4 * g + 2] = x.z; o[4 * g + 3] = x.w;
        }
    } else if constexpr (KPL == 2) {
        float2 x = *reinterpret_cast<const float2*>(base + lane * 2);
        o[0] = x.x; o[1] = x.y;
    } else {
        o[0] = base[lane];
    }
}

template <int KPL, typename V>
__device__ __forceinline__ void store_kvec(int32_t* __restrict__ base, int lane, const V& o) {
    if constexpr (KPL >= 4) {
#pragma unroll
        for (int g = 0; g < KPL / 4; ++g)
            *reinterpret_cast<int4*>(base + g * 256 + lane * 4) = make_int4(o[4 * g], o[4 * g + 1], o[4 * g + 2], o[4 * g + 3]);
    } else if constexpr (KPL == 2) {
        *reinterpret_cast<int2*>(base + lane * 2) = make_int2(o[0], o[1]);
    } else {
        base[lane] = o[0];
    }
}

// A K-vector of u16 cells in the lane layout (same positions as load_kvec):
// for KPL >= 4 each lane reads 8 bytes per group of 4 topics, 512 contiguous
// bytes per wavefront load.
template <int KPL, typename V>
__device__ __forceinline__ void load_kvec_u16(const uint16_t* __restrict__ base, int lane, V& o) {
    if constexpr (KPL >= 4) {
#pragma unroll
        for (int g = 0; g < KPL / 4; ++g) {
            const uint2 x = *reinterpret_cast<const uint2*>(base + g * 256 + lane * 4);
            o[4 * g] = (int)(x.x & 0xFFFFu); o[4 * g + 1] = (int)(x.x >> 16);
            o[4 * g + 2] = (int)(x.y & 0xFFFFu); o[4 * g + 3] = (int)(x.y >> 16);
        }
    } else if constexpr (KPL == 2) {
        const uint32_t x = *reinterpret_cast<const uint32_t*>(base + lane * 2);
        o[0] = (int)(x & 0xFFFFu); o[1] = (int)(x >> 16);
    } else {
        o[0] = (int)base[lane];
    }
}

// ---- cross-lane scan: DPP row_shr 1,2,4,8 then row_bcast 15 / 31 ----
__device__ __forceinline__ float dpp_row_shr1(float v) { return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xF, 0xF, true)); }
__device__ __forceinline__ float dpp_row_shr2(float v) { return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x112, 0xF, 0xF, true)); }
__device__ __forceinline__ float dpp_row_shr4(float v) { return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x114, 0xF, 0xF, true)); }
__device__ __forceinline__ float dpp_row_shr8(float v) { return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x118, 0xF, 0xF, true)); }
__device__ __forceinline__ float dpp_bcast15(float v) { return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x142, 0xA, 0xF, false)); }
__device__ __forceinline__ float dpp_bcast31(float v) { return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x143, 0xC, 0xF, false)); }
__device__ __forceinline__ float dpp_wave_shr1(float v) { return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, true)); }

// Inclusive scan over the 64 lanes in the order the CPU oracle emulates
// (oracle/lda_oracle.cpp prefix_sum_device).
__device__ __forceinline__ float wave_inclusive_scan(float v) {
    v = v + dpp_row_shr1(v);
    v = v + dpp_row_shr2(v);
    v = v + dpp_row_shr4(v);
    v = v + dpp_row_shr8(v);
    v = v + dpp_bcast15(v);
    v = v + dpp_bcast31(v);
    return v;
}

__device__ __forceinline__ int readlane_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float readlane_f(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }

template <int N> using ivec = int __attribute__((ext_vector_type(N)));
template <int N> using fvec = float __attribute__((ext_vector_type(N)));

// Global-address-space views.  Pointers read from device memory (the fold-in
// parameter block) are generic to the compiler, and generic (flat) accesses
// make it wait for every outstanding load and LDS operation at each use;
// these views turn them into global loads that are tracked one by one.
template <typename T> using gbl = __attribute__((address_space(1))) T;
template <typename T>
__device__ __forceinline__ gbl<T>* gptr(T* p) { return (gbl<T>*)p; }

template <int KPL, typename V>
__device__ __forceinline__ void gload_kvec(const gbl<const int32_t>* base, int lane, V& o) {
    if constexpr (KPL >= 4) {
#pragma unroll
        for (int g = 0; g < KPL / 4; ++g) {
            const ivec<4> x = *reinterpret_cast<const gbl<const ivec<4>>*>(base + g * 256 + lane * 4);
            o[4 * g] = x[0]; o[4 * g + 1] = x[1]; o[4 * g + 2] = x[2]; o[4 * g + 3] = x[3];
        }
    } else if constexpr (KPL == 2) {
        const ivec<2> x = *reinterpret_cast<const gbl<const ivec<2>>*>(base + lane * 2);
        o[0] = x[0]; o[1] = x[1];
    } else {
        o[0] = base[lane];
    }
}

template <int KPL>
constexpr int kpl_shift() {
    return (KPL == 1) ? 0 : (KPL == 2) ? 1 : (KPL == 4) ? 2 : (KPL == 8) ? 3 : (KPL == 16) ? 4 : 5;
}


template <int KPL>
__device__ __forceinline__ void lane_limits(int K, int lane, int& nvalid, int& lastLane) {
    int nv = K - lane * KPL;
    nvalid = nv < 0 ? 0 : (nv > KPL ? KPL : nv);
    lastLane = (K - 1) / KPL;
}

// The number of a lane's cdf values cdf_j = excl + w[j] (j < KPL) that are
// <= r.  w is non-decreasing (a running sum of non-negative weights) and
// rounding is monotone, so the cdf values are too and {j : cdf_j <= r} is a
// prefix: a binary search finds its length from log2(KPL) probes, each probe
// chosen among the candidates by the earlier decisions (a select tree, since
// the index differs per lane), plus the last value, which only counts when
// every probe did.  Equal to counting every j.
template <int KPL, int BASE, int STRIDE, int LEVELS, int B0>
__device__ __forceinline__ float pick_probe(const fvec<KPL>& w, const bool* b) {
    if constexpr (LEVELS == 0) {
        return w[BASE];
    } else {
        const float lo = pick_probe<KPL, BASE, STRIDE, LEVELS - 1, B0 + 1>(w, b);
        const float hi = pick_probe<KPL, BASE + STRIDE * (1 << (LEVELS - 1)), STRIDE, LEVELS - 1, B0 + 1>(w, b);
        return b[B0] ? hi : lo;
    }
}

template <int KPL, int STEP>
__device__ __forceinline__ void probe_steps(const fvec<KPL>& w, float excl, float r, bool* b) {
    if constexpr ((1 << STEP) < KPL) {
        // probe index (decisions so far) * KPL/2^STEP + KPL/2^(STEP+1) - 1
        const float c = pick_probe<KPL, KPL / (2 << STEP) - 1, KPL / (1 << STEP), STEP, 0>(w, b);
        b[STEP] = !(r < excl + c);
        probe_steps<KPL, STEP + 1>(w, excl, r, b);
    }
}

template <int KPL>
__device__ __forceinline__ int cdf_count_le(const fvec<KPL>& w, float excl, float r, float last) {
    bool b[6] = {false, false, false, false, false, false};
    probe_steps<KPL, 0>(w, excl, r, b);
    int n = (r < last) ? 0 : 1;
#pragma unroll
    for (int s = 0; (1 << s) < KPL; ++s) n += b[s] ? (KPL >> (s + 1)) : 0;
    return n;
}

// Inclusive scan of the lane-layout weights (lane-local, then the DPP wave
// scan) and the inverse-CDF search of sampleFromDiscreteAcc (sample.hpp:133-189):
// the first topic with u * total < cdf, else K-1.  Returns the topic
// (wave-uniform); tot receives cdf[K-1].
template <int KPL>
__device__ __forceinline__ int scan_search(fvec<KPL>& w, float u, int K, int nvalid, int lastLane, float& tot) {
#pragma unroll
    for (int j = 1; j < KPL; ++j) w[j] = w[j - 1] + w[j];
    float incl = wave_inclusive_scan(w[KPL - 1]);
    float excl = dpp_wave_shr1(incl);
    // the lane's cdf values are cdf_j = excl + w[j] (float adds, as the scan
    // defines them); only the probed ones and the last are formed
    const float last = excl + w[KPL - 1];
    tot = readlane_f(last, lastLane);
    const float r = u * tot;
    const int miss = cdf_count_le<KPL>(w, excl, r, last);
    const unsigned long long m = __ballot(miss < nvalid);
    int zn;
    if (m == 0ull) zn = K - 1;
    else {
        const int L = __builtin_ctzll(m);
        zn = L * KPL + readlane_i(miss, L);
    }
    return uniform_i(zn);
}

constexpr int MODE_DET = 0;     // exclusive cell owner: plain stores to n_kw
constexpr int MODE_ATOMIC = 1;  // shared n_kw: exact int32 atomics

// forShuffled's prime table (Utils.hpp:159-163).
__constant__ uint64_t kShufflePrimes[16] = {65537, 65539, 65543, 65551, 65557, 65563, 65579, 65581,
                                            65587, 65599, 65609, 65617, 65629, 65633, 65647, 65651};

// tomoto's max(val - w, 0): std::max(a, 0) keeps a unless a < 0
__device__ __forceinline__ float sub_clamp(float v, float w) {
    const float a = v - w;
    return a < 0.0f ? 0.0f : a;
}

// max(x - w, 0) on a shared float with a compare-and-swap loop
__device__ __forceinline__ void atomic_sub_clamp(float* p, float w) {
    unsigned int* u = reinterpret_cast<unsigned int*>(p);
    unsigned int old = *u, assumed;
    do {
        assumed = old;
        const float nv = sub_clamp(__uint_as_float(assumed), w);
        old = atomicCAS(u, assumed, __float_as_uint(nv));
    } while (old != assumed);
}

// Lane-layout position of topic k for a runtime KPL (kpos).
__device__ __forceinline__ int kpos_rt(int kpl, int k) {
    const int l = k / kpl, j = k - l * kpl;
    return kpl < 4 ? k : (j >> 2) * 256 + l * 4 + (j & 3);
}

// With word priors the initial topic is drawn from the word's eta column by
// sampleFromDiscrete (sample.hpp:121-131, LDAModel.hpp:769-773): r = u * (the
// sequential float sum of the column), then subtract eta_kv while r > eta_kv
// and k < K-1.  A word without a prior has the column eta everywhere.
__device__ __forceinline__ int draw_from_eta_column(const float* col, int kpl, float eta, int K, float u) {
    float tot = 0.0f;
    for (int k = 0; k < K; ++k) tot += col ? col[kpos_rt(kpl, k)] : eta;
    float r = u * tot;
    int k = 0;
    for (;;) {
        const float w = col ? col[kpos_rt(kpl, k)] : eta;
        if (!(r > w && k < K - 1)) break;
        r -= w;
        ++k;
    }
    return k;
}

__device__ __forceinline__ double wave_sum_d(double x) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

}  // namespace
}  // namespace lda
