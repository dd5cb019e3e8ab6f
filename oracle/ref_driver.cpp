// ref_driver.cpp -- extern "C" entry points over the reference's own
// header-only utilities, compiled in place from /root/reference/src/Utils by
// oracle/Makefile (target `ref`) into oracle/_ref/.  TEST INFRASTRUCTURE ONLY:
// used by tests/ to pin the oracle's restatements bit-for-bit.  No reference
// source is copied; this file only calls it.
#include <cassert>
#include <cstdint>
#include <utility>
#include <vector>

#include "exception.h"  // exc:: types used by AliasMethod.hpp (included by sample.hpp)
#include "sample.hpp"   // tomoto::sample::prefixSum, sampleFromDiscreteAcc
#include "math.h"       // tomoto::math::lgammaT, digammaT (LUT.hpp)
#include "Utils.hpp"    // tomoto::forShuffled, sortAndWriteOrder

namespace {
// Adapter: sampleFromDiscreteAcc draws one rg.uniform_real(); this returns a
// fixed value so the search can be driven with chosen uniforms.
struct FixedUniform {
    float u;
    float uniform_real() { return u; }
};
}  // namespace

extern "C" {

void ref_prefix_sum(float* a, int n) { tomoto::sample::prefixSum(a, n); }

int ref_sample_acc(const float* cdf, int K, float u) {
    FixedUniform rg{u};
    return (int)tomoto::sample::sampleFromDiscreteAcc(cdf, cdf + K, rg);
}

float ref_lgammaT(float x) { return tomoto::math::lgammaT(x); }
float ref_digammaT(float x) { return tomoto::math::digammaT(x); }

void ref_for_shuffled(uint64_t N, uint64_t seed, uint32_t* out) {
    size_t i = 0;
    tomoto::forShuffled((size_t)N, (size_t)seed, [&](size_t id) { out[i++] = (uint32_t)id; });
}

// sortAndWriteOrder over (cf, df) pairs with removeStopwords' ordering rule
// (TopicModel.hpp:539-551: filtered words after kept ones, each group by
// (cf, df) descending) and the rm_top rotation.
void ref_vocab_order(uint64_t n, const uint64_t* cf, const uint64_t* df, uint64_t min_cf, uint64_t min_df,
                     uint64_t rm_top, uint32_t* order_out) {
    std::vector<std::pair<size_t, size_t>> v(n);
    for (uint64_t i = 0; i < n; ++i) v[i] = {cf[i], df[i]};
    std::vector<uint32_t> order;
    auto bad = [&](const std::pair<size_t, size_t>& a) { return a.first < min_cf || a.second < min_df; };
    tomoto::sortAndWriteOrder(v, order, rm_top, [&](const std::pair<size_t, size_t>& a, const std::pair<size_t, size_t>& b) {
        if (bad(a)) return bad(b) ? a > b : false;
        if (bad(b)) return true;
        return a > b;
    });
    for (uint64_t i = 0; i < n; ++i) order_out[i] = order[i];
}
}
