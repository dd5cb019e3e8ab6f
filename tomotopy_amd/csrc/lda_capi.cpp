// lda_capi.cpp -- host core and C ABI of the MI355X LDA engine.
//
// The host core plays the part of tomoto's TopicModel / LDAModel driver for
// plain LDA (src/TopicModel/TopicModel.hpp:684-755, LDAModel.hpp:356-674):
// it owns the device state of one model, resolves the parallel scheme,
// computes the partition schedule, runs the iteration loop with the alpha
// fixed point, and reduces the FP64 log-likelihood partials.  The per-token
// work runs in the kernels of lda_device.hip.  Nothing here falls back to a
// CPU sampler: without a HIP device every call fails with LDA_ERR_DEVICE.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "lda_host.h"

namespace lda_host {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

hipError_t FoldStreams::create() {
    for (int b = 0; b < lda::kFoldBuckets - 1; ++b) {
        hipError_t e = hipStreamCreateWithFlags(&aux[b], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&join[b], hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    return hipEventCreateWithFlags(&fork, hipEventDisableTiming);
}

void FoldStreams::destroy() {
    for (int b = 0; b < lda::kFoldBuckets - 1; ++b) {
        if (aux[b]) (void)hipStreamDestroy(aux[b]);
        if (join[b]) (void)hipEventDestroy(join[b]);
        aux[b] = nullptr;
        join[b] = nullptr;
    }
    if (fork) (void)hipEventDestroy(fork);
    fork = nullptr;
}

hipError_t launch_fold_buckets(int kpl, bool asym, const lda::InferParams* p_dev, const int grids[lda::kFoldBuckets],
                               hipStream_t s, FoldStreams& fs) {
    hipError_t e = hipEventRecord(fs.fork, s);
    // the longest documents (most segments) first
    for (int b = lda::kFoldBuckets - 1; b >= 1 && e == hipSuccess; --b) {
        if (grids[b] <= 0) continue;
        e = hipStreamWaitEvent(fs.aux[b - 1], fs.fork, 0);
        if (e == hipSuccess) e = lda::launch_infer(kpl, 1 << b, asym, p_dev, grids[b], fs.aux[b - 1]);
        if (e == hipSuccess) e = hipEventRecord(fs.join[b - 1], fs.aux[b - 1]);
    }
    if (e == hipSuccess && grids[0] > 0) e = lda::launch_infer(kpl, 1, asym, p_dev, grids[0], s);
    for (int b = 1; b < lda::kFoldBuckets && e == hipSuccess; ++b)
        if (grids[b] > 0) e = hipStreamWaitEvent(s, fs.join[b - 1], 0);
    return e;
}

}  // namespace lda_host

using lda_host::fail;
using lda_host::kMaxAlpha;
using lda_host::kMaxDocTokens;

// f(lo, hi) over [0, n) in contiguous ranges on up to 16 host threads (one
// range per grain items at least), the calling thread taking the first.
template <class F>
static void host_parallel(int64_t n, int64_t grain, const F& f) {
    const int64_t hc = std::max<int64_t>(1, (int64_t)std::thread::hardware_concurrency());
    const int64_t nt = std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(hc, 16), (n + grain - 1) / grain));
    if (nt == 1) {
        f((int64_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve((size_t)nt - 1);
    for (int64_t t = 1; t < nt; ++t) th.emplace_back([&f, n, nt, t] { f(n * t / nt, n * (t + 1) / nt); });
    f((int64_t)0, n / nt);
    for (auto& x : th) x.join();
}

// Lowers `at` to v if v is smaller (the first failing index over threads).
static void atomic_min_i64(std::atomic<int64_t>& at, int64_t v) {
    int64_t cur = at.load();
    while (v < cur && !at.compare_exchange_weak(cur, v)) {}
}

namespace {


// tomoto's float LUT digamma (src/Utils/LUT.hpp:10-79, math.h:134-154), used
// for dAlpha in the alpha fixed point (LDAModel.hpp:224).  Grid values are a
// double-precision digamma rounded to float; steps 1/1024, 1/64, 1/4.
double digamma_d(double x) {
    double r = 0.0;
    while (x < 6.0) {
        r -= 1.0 / x;
        x += 1.0;
    }
    double f = 1.0 / (x * x);
    double t = f * (-1.0 / 12 + f * (1.0 / 120 + f * (-1.0 / 252 + f * (1.0 / 240 + f * (-1.0 / 132)))));
    return r + std::log(x) - 0.5 / x + t;
}

struct DigammaLut {
    static constexpr int N = 1024, M = 6400, L = 4000;
    const float P = 1.0f / 1024, Q = 1.0f / 64, R = 1.0f / 4;
    std::vector<float> pts;
    DigammaLut() : pts(N + M + L) {
        for (int i = 0; i < N; ++i) pts[i] = (float)digamma_d(i ? (double)(i * P) : (double)0.0001f);
        for (int i = 0; i < M; ++i) pts[N + i] = (float)digamma_d((double)(i * Q + N * P));
        for (int i = 0; i < L; ++i) pts[N + M + i] = (float)digamma_d((double)(i * R + N * P + M * Q));
    }
    float operator()(float x) const {
        if (!std::isfinite(x)) return (std::isnan(x) || x < 0) ? NAN : INFINITY;
        if (x < 0) return NAN;
        if (x < 0.001f) {
            if (x == 0) return -INFINITY;
            return logf(x + 2) - 0.5f / (x + 2) - 1 / 12.f / powf(x + 2, 2) - 1 / (x + 1) - 1 / x;
        }
        if (x >= N * P + M * Q + (L - 1) * R) return logf(x) - 0.5f / x - 1 / 12.f / powf(x, 2);
        size_t idx;
        float a, nx = x;
        if (x < N * P) {
            idx = (size_t)(nx / P);
            a = (nx - idx * P) / P;
        } else {
            nx -= N * P;
            if (nx < M * Q) {
                idx = (size_t)(nx / Q);
                a = (nx - idx * Q) / Q;
                idx += N;
            } else {
                nx -= M * Q;
                idx = (size_t)(nx / R);
                a = (nx - idx * R) / R;
                idx += N + M;
            }
        }
        return pts[idx] + a * (pts[idx + 1] - pts[idx]);
    }
};

const DigammaLut& digamma_lut() {
    static DigammaLut t;
    return t;
}

int kpl_for(int K) {
    int kpl = 1;
    while (64 * kpl < K) kpl <<= 1;
    return kpl;
}

}  // namespace


namespace {

// Persistent grids of the fold-in kernel per segment bucket, and the largest
// segment count that fits, for the context's variant (word priors or not).
// Drops a cached fold-in graph, which launches the other variant.
void compute_fold_grid(lda_ctx* c) {
    c->seg_supported = 1;
    for (int b = 0; b < lda::kFoldBuckets; ++b) c->fold_grid[b] = 0;
    for (int b = 0; b < lda::kFoldBuckets; ++b) {
        const int per_cu = lda::infer_blocks_per_cu(c->kpl, 1 << b, c->has_priors);
        c->fold_grid[b] = c->num_cu * per_cu;
        if (per_cu > 0) c->seg_supported = 1 << b;
        else break;
    }
    if (c->seg_max > c->seg_supported) c->seg_max = c->seg_supported;
    if (c->infer_exec) {
        (void)hipGraphExecDestroy(c->infer_exec);
        c->infer_exec = nullptr;
    }
}

int upload_alpha(lda_ctx* c) {
    std::vector<float> a(c->Kp, 0.0f);
    for (int k = 0; k < c->K; ++k) a[c->pos(k)] = c->alpha[k];
    std::vector<float> ca(c->K);
    float s = 0.0f;
    for (int k = 0; k < c->K; ++k) {
        s = s + c->alpha[k];
        ca[k] = s;
    }
    HIP_TRY(hipMemcpyAsync(c->d_alpha.p, a.data(), a.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_alpha_cdf.p, ca.data(), ca.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LDA_OK;
}

// Vocabulary chunks of the partition scheme at cumulative collection
// frequency quantiles (LDAModel.hpp:436-449), padded with V when the
// quantiles produce fewer than W boundaries; then the per-document token
// offset of every chunk (LDAModel.hpp:451-469; words are sorted per document).
int prepare_partition(lda_ctx* c, int W) {
    if (c->part_W == W && c->d_coff.p) return LDA_OK;
    std::vector<uint64_t> cf(c->V, 0);
    if (!c->chunk_cf.empty()) cf = c->chunk_cf;
    else
        for (uint32_t v : c->words)
            if (v < (uint32_t)c->V) ++cf[v];
    uint64_t tot = 0;
    for (auto x : cf) tot += x;
    std::vector<uint32_t> ch;
    uint64_t cum = 0;
    for (int i = 0; i < c->V; ++i) {
        cum += cf[i];
        if (cum * (uint64_t)W >= tot * (uint64_t)(ch.size() + 1)) ch.push_back((uint32_t)(i + 1));
    }
    while ((int)ch.size() < W) ch.push_back((uint32_t)c->V);
    ch.resize(W);
    ch[W - 1] = (uint32_t)c->V;
    const int64_t D = c->D;
    std::vector<uint64_t> coff((size_t)(W + 1) * D);
    for (int64_t d = 0; d < D; ++d) {
        uint64_t b = c->off[d], e = c->off[d + 1];
        coff[d] = b;
        for (int p = 0; p < W; ++p) {
            auto it = std::lower_bound(c->words.begin() + b, c->words.begin() + e, ch[p]);
            coff[(size_t)(p + 1) * D + d] = (uint64_t)(it - c->words.begin());
        }
        coff[(size_t)W * D + d] = e;
    }
    HIP_TRY(c->d_coff.alloc(coff.size()));
    HIP_TRY(hipMemcpyAsync(c->d_coff.p, coff.data(), coff.size() * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c->d_nk_priv.alloc((size_t)W * c->Kp));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->part_W = W;
    return LDA_OK;
}

lda::SweepParams base_params(lda_ctx* c) {
    lda::SweepParams p{};
    p.words = c->d_words.p;
    p.z = c->d_z.p;
    p.off = c->d_off.p;
    p.ndk = c->d_ndk.p;
    p.nkw = c->d_nkw.p;
    p.alpha = c->d_alpha.p;
    p.eta = c->eta;
    p.veta = c->veta;
    p.K = c->K;
    p.V = c->V;
    p.Kp = c->Kp;
    p.seed_lo = (uint32_t)c->seed;
    p.seed_hi = (uint32_t)(c->seed >> 32);
    p.sweep = (uint32_t)c->global_step;
    p.D = c->D;
    if (c->has_priors) {
        p.prior_row = c->d_prior_row.p;
        p.prior_eta = c->d_prior_eta.p;
        p.etasum = c->d_etasum.p;
    }
    if (c->narrow_active) {
        p.nkw16 = c->d_nkw16.p;
        p.nkw_hot = c->nkw_hot;
    }
    p.exact = c->exact_counts() ? 1 : 0;
    return p;
}

// Narrow columns for the device-wide sweeps of one lda_train / lda_sweep
// call (integer counts): pack rows [nkw_hot, V) of n_kw into u16 once at the
// start of the call and write them back at its end, so that every other
// kernel keeps reading int32 columns.  When a cell is outside [0, 2^15) (counts
// set from outside that do not match the corpus) the call runs on int32
// columns instead.
int narrow_begin(lda_ctx* c, int W, bool copy_merge) {
    c->narrow_active = false;
    c->narrow_used = false;
    if (!c->narrow_enabled || W != 0 || copy_merge || c->tw != LDA_TW_ONE || c->nkw_hot >= c->V ||
        c->kpl > lda::kSmallMaxKpl || c->exact_counts())
        return LDA_OK;
    const hipStream_t s = c->stream;
    HIP_TRY(c->d_nkw16.alloc((size_t)(c->V - c->nkw_hot) * c->Kp));
    HIP_TRY(c->d_bad.alloc(1));
    HIP_TRY(lda::launch_fill_u32(c->d_bad.p, 0u, 1, s));
    HIP_TRY(lda::launch_pack_narrow(c->d_nkw.p, c->d_nkw16.p, c->nkw_hot, c->V, c->Kp, c->d_bad.p, s));
    uint32_t bad = 0;
    HIP_TRY(hipMemcpyAsync(&bad, c->d_bad.p, sizeof bad, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    c->narrow_active = bad == 0;
    c->narrow_used = c->narrow_active;
    return LDA_OK;
}

int narrow_end(lda_ctx* c) {
    if (!c->narrow_active) return LDA_OK;
    c->narrow_active = false;
    HIP_TRY(lda::launch_unpack_narrow(c->d_nkw.p, c->d_nkw16.p, c->nkw_hot, c->V, c->Kp, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LDA_OK;
}

// Worker counts the device picks when the caller passes workers == 0 with an
// explicit PARTITION or COPY_MERGE (tomotopy passes 0 for "as many as the
// hardware has", TopicModel.hpp:686).
//   partition: each round gives every worker about n_valid / W^2 tokens; W is
//   the largest count that keeps that at kPartitionRoundTokens or more (a
//   round shorter than that is mostly launch and n_dk staging), within
//   tomoto's cap max((V+3)/4, 1) (LDAModel.hpp:491-503), the documents and
//   the resident capacity of the round kernel.
//   copy_merge: tomoto's cap max((D+1)/2, 1), at most 64 workers and at most
//   32 GB of private n_kw copies (a ninth of the HBM).
constexpr double kPartitionRoundTokens = 48.0;

int device_partition_workers(lda_ctx* c) {
    const int cap = std::max((c->V + 3) / 4, 1);
    const int resident = c->num_cu * lda::sweep_det_blocks_per_cu(c->kpl, c->has_priors, c->tw != LDA_TW_ONE);
    int W = (int)std::floor(std::sqrt((double)c->n_valid / kPartitionRoundTokens));
    W = std::min(W, cap);
    W = std::min<int64_t>(W, std::max<int64_t>(c->D, 1));
    W = std::min(W, std::min(resident, 1024));
    return std::max(W, 1);
}

int device_copy_workers(lda_ctx* c) {
    const int64_t cap_cm = std::max<int64_t>((c->D + 1) / 2, 1);
    const double per_copy = std::max(1.0, (double)c->V * c->Kp * 4.0);
    const int64_t by_mem = std::max<int64_t>((int64_t)(32e9 / per_copy), 1);
    return (int)std::min<int64_t>(std::min<int64_t>(cap_cm, by_mem), 64);
}

// Concurrency of the device-wide schedule.  Every document in flight samples
// against columns and n_k that miss the other in-flight documents' changes
// (AD-LDA staleness); the larger the share of the corpus in flight, the more
// the chain lags tomoto's partition chain while it is still descending (on
// the K=1000 corpus cut to 10,000 documents, 41% in flight lagged by up to
// 0.66% of ll_per_word, an eighth by at most 0.13%).  So by default at most
// ceil(D / kInFlightDiv) documents are in flight: as 1-wavefront workgroups
// (one per CU where possible) when that is below the resident capacity of
// those, else as 8-wavefront workgroups up to the resident capacity.  Large
// corpora (D >= 8 x resident wavefronts) run the full grid.
constexpr int64_t kInFlightDiv = 8;
constexpr int kBigSweepWaves = 4;


// Tokens per group of shared n_kw atomics in the device-wide sweeps (K <=
// 2048).  A wavefront sends its tokens' n_kw changes once per group, as whole-
// wavefront atomics; per-token single-lane atomics held every later load wait
// behind an atomic's round trip (K = 1000, 30M tokens, 4,096 wavefronts: 14.0
// ms per sweep; groups of 4 / 8 / 16 / 32: 12.1 / 11.8 / 11.5 / 11.1 ms).  The
// changes held back meanwhile are one more source of staleness: on the 3M-token
// cut of that corpus (1,250 wavefronts) the chain's largest lag behind the
// oracle's partition schedule while descending grows from 0.15% of
// ll_per_word (groups of one token) to 0.17 / 0.21 / 0.25 / 0.34% for those
// groups.  The group is the largest power of two up to 32 that keeps the
// changes held back by all wavefronts at most 1/128 of the tokens.
constexpr int64_t kHeldBackDiv = 128;

int device_col_group(const lda_ctx* c, int64_t wavefronts) {
    if (c->dev_col_group > 0) return c->dev_col_group;
    int g = 32;
    while (g > 1 && (int64_t)g * wavefronts * kHeldBackDiv > (int64_t)c->n_valid) g >>= 1;
    return g;
}

int device_wide_grid(lda_ctx* c, int* waves_out, int* rows_out = nullptr) {
    if (rows_out) *rows_out = 0;
    const bool wtd = c->tw != LDA_TW_ONE;
    auto per_cu = [&](int waves) {
        return wtd ? lda::sweep_device_w_blocks_per_cu(c->kpl, c->has_priors, waves)
                   : lda::sweep_device_blocks_per_cu(c->kpl, c->has_priors, waves);
    };
    if (!c->dev_auto) {
        const bool one = c->kpl > lda::kSmallMaxKpl || c->exact_counts();
        const int waves = one ? 1 : c->dev_waves;
        int grid = c->num_cu * per_cu(waves);
        if (c->dev_workers > 0) grid = std::min(grid, c->dev_workers);
        *waves_out = waves;
        return grid;
    }
    const int64_t inflight = std::max<int64_t>((c->D + kInFlightDiv - 1) / kInFlightDiv, 1);
    if (c->kpl > lda::kSmallMaxKpl || c->exact_counts()) {
        // K above 2048 or exact counts.  While at least two wavefronts fit a
        // workgroup with alpha, the shared n_k and each wavefront's n_dk row
        // and column copy in LDS (K <= 4096: three), that workgroup, so that a
        // token reads no device memory (K = 4096, 3M tokens: 18.7 ms per sweep
        // against 22.1 with the rows in device memory; with only n_dk in LDS
        // 19.3 ms at six wavefronts, 21.0 at four; at K = 8192 one such
        // wavefront per CU takes 95 ms, two with only n_dk 64 ms, against
        // 56.5 ms for the one-wavefront workers below).  Otherwise workgroups
        // of kBigSweepWaves wavefronts sharing alpha and n_k in LDS while four
        // such workgroups fit a CU, else one-wavefront workers with
        // device-memory n_k.
        if (c->kpl > lda::kSmallMaxKpl) {
            for (int w = kBigSweepWaves; w >= 2; --w) {
                const int per = lda::bk_sweep_device_lds_blocks_per_cu(wtd, c->has_priors, c->kpl, c->Kp, w, 3);
                if (per <= 0) continue;
                *waves_out = w;
                if (rows_out) *rows_out = 3;
                const int64_t wg = (inflight + w - 1) / w;
                return (int)std::max<int64_t>(1, std::min<int64_t>(wg, (int64_t)c->num_cu * per));
            }
        }
        if (lda::bk_sweep_device_lds_bytes(c->kpl, c->Kp, kBigSweepWaves, 0) <= 40 * 1024) {
            const int per =
                lda::bk_sweep_device_lds_blocks_per_cu(wtd, c->has_priors, c->kpl, c->Kp, kBigSweepWaves, 0);
            if (per > 0) {
                *waves_out = kBigSweepWaves;
                const int64_t wg = (inflight + kBigSweepWaves - 1) / kBigSweepWaves;
                return (int)std::max<int64_t>(1, std::min<int64_t>(wg, (int64_t)c->num_cu * per));
            }
        }
        *waves_out = 1;
        const int per = std::max(lda::bk_sweep_blocks_per_cu(wtd, c->has_priors, true, c->kpl), 1);
        return (int)std::min<int64_t>(inflight, (int64_t)c->num_cu * per);
    }
    const int W8 = lda::kMaxSweepWaves;
    const int full8 = c->num_cu * per_cu(W8);
    if (inflight >= (int64_t)full8 * W8) {
        *waves_out = W8;
        return full8;
    }
    const int full1 = c->num_cu * per_cu(1);
    if (inflight <= full1) {
        *waves_out = 1;
        return (int)inflight;
    }
    *waves_out = W8;
    return (int)std::min<int64_t>((inflight + W8 - 1) / W8, full8);
}

// Scheme resolution as in TopicModel::train (TopicModel.hpp:684-690):
// workers == 1 or NONE -> one worker; the partition worker count is capped at
// max((V+3)/4, 1) and the copy_merge one at max((D+1)/2, 1)
// (LDAModel.hpp:491-503).  workers == 0 with PARTITION or COPY_MERGE runs
// that scheme with the device's worker count; workers == 0 with DEFAULT
// selects the device-wide schedule (every resident workgroup a worker).
// Returns the worker count (0 = device-wide) and whether copy_merge applies.
int resolve_workers(lda_ctx* c, int workers, int scheme, int* W_out, bool* copy_merge) {
    *copy_merge = false;
    if (workers < 0) return fail(LDA_ERR_INVALID_ARGUMENT, "workers must be >= 0");
    if (scheme < LDA_PS_DEFAULT || scheme > LDA_PS_PARTITION) return fail(LDA_ERR_INVALID_ARGUMENT, "invalid ParallelScheme");
    if (scheme == LDA_PS_NONE || workers == 1) {
        *W_out = 1;
        return LDA_OK;
    }
    if (workers == 0 && scheme == LDA_PS_DEFAULT) {
        *W_out = 0;
        return LDA_OK;
    }
    if (scheme == LDA_PS_COPY_MERGE) {
        const int64_t cap_cm = std::max<int64_t>((c->D + 1) / 2, 1);
        int W = workers == 0 ? device_copy_workers(c) : (int)std::min<int64_t>(std::min<int64_t>(workers, cap_cm), 64);
        if (W <= 1) { *W_out = 1; return LDA_OK; }
        const double bytes = (double)W * c->V * c->Kp * 4.0;
        if (bytes > 64e9) return fail(LDA_ERR_OUT_OF_MEMORY, "copy_merge with %d workers needs %.1f GB of n_kw copies", W, bytes / 1e9);
        *W_out = W;
        *copy_merge = true;
        return LDA_OK;
    }
    int W;
    if (workers == 0) W = device_partition_workers(c);
    else W = std::min(std::min(workers, std::max((c->V + 3) / 4, 1)), 1024);
    *W_out = W;
    return LDA_OK;
}

// The W rounds of one partition sweep (LDAModel.hpp:373-399), captured once
// as a graph of W launches and replayed; the sweep number, which keys the
// rounds' random streams, is written to d_sweep by a fill kernel first.  The
// graph is rebuilt when the buffers or sizes it was captured with change.
bool same_round_params(const lda::SweepParams& a, const lda::SweepParams& b) {
    return a.words == b.words && a.z == b.z && a.off == b.off && a.ndk == b.ndk && a.nkw == b.nkw &&
           a.nk_in == b.nk_in && a.nk_out == b.nk_out && a.alpha == b.alpha && a.eta == b.eta && a.veta == b.veta &&
           a.K == b.K && a.V == b.V && a.Kp == b.Kp && a.seed_lo == b.seed_lo && a.seed_hi == b.seed_hi &&
           a.D == b.D && a.coff == b.coff && a.W == b.W && a.wt == b.wt && a.prior_row == b.prior_row &&
           a.prior_eta == b.prior_eta && a.etasum == b.etasum && a.sweep_dev == b.sweep_dev;
}

int run_partition_rounds(lda_ctx* c, lda::SweepParams p, int W, bool weighted) {
    const hipStream_t s = c->stream;
    HIP_TRY(c->d_sweep.alloc(1));
    p.sweep_dev = c->d_sweep.p;
    const lda::SweepParams key = p;
    const bool same = c->part_exec && c->part_key_w == weighted && same_round_params(key, c->part_key);
    if (!same) {
        if (c->part_exec) {
            (void)hipGraphExecDestroy(c->part_exec);
            c->part_exec = nullptr;
        }
        hipGraph_t g = nullptr;
        HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipSuccess;
        for (int r = 0; r < W && e == hipSuccess; ++r) {
            p.round = r;
            e = weighted ? lda::launch_sweep_det_w(c->kpl, p, s) : lda::launch_sweep_det(c->kpl, p, s);
        }
        const hipError_t e2 = hipStreamEndCapture(s, &g);
        if (e == hipSuccess) e = e2;
        if (e == hipSuccess) e = hipGraphInstantiate(&c->part_exec, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            c->part_exec = nullptr;
            return fail(LDA_ERR_DEVICE, "partition graph: %s", hipGetErrorString(e));
        }
        c->part_key = key;
        c->part_key_w = weighted;
    }
    HIP_TRY(lda::launch_fill_u32(c->d_sweep.p, p.sweep, 1, s));
    HIP_TRY(hipGraphLaunch(c->part_exec, s));
    return LDA_OK;
}

// The time of the device-wide sweep whose events are pending (see lda_ctx).
int dw_read(lda_ctx* c) {
    if (c->dw_pending < 0) return LDA_OK;
    const int k = c->dw_pending;
    c->dw_pending = -1;
    HIP_TRY(hipEventSynchronize(c->dw_ev[k][1]));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, c->dw_ev[k][0], c->dw_ev[k][1]));
    c->stat_ms += ms;
    return LDA_OK;
}

int sweep_once(lda_ctx* c, int W, bool copy_merge) {
    lda::SweepParams p = base_params(c);
    const hipStream_t s = c->stream;
    if (c->tw != LDA_TW_ONE) {
        // term-weighted: device-wide, copy_merge or partition, then tomoto's
        // weighted merge n_kw = max(n_kw, 0), n_k = rowsum (LDAModel.hpp:565-579)
        p.wt = c->d_wt.p;
        float ms = 0.0f;
        int dw = -1;  // the device-wide sweep's event pair
        if (copy_merge) {
            const int64_t n = (int64_t)c->V * c->Kp;
            HIP_TRY(c->d_nkw_loc.alloc((size_t)W * std::max<int64_t>(n, 1)));
            HIP_TRY(c->d_nk_priv.alloc((size_t)W * c->Kp));
            c->part_W = 0;  // d_nk_priv now holds copy_merge state
            HIP_TRY(lda::launch_replicate(c->d_nkw_loc.p, c->d_nkw.p, n, W, s));
            HIP_TRY(lda::launch_replicate(c->d_nk_priv.p, c->d_nk.p, c->Kp, W, s));
            p.nkw = c->d_nkw_loc.p;
            p.nk_in = c->d_nk_priv.p;
            p.nk_out = c->d_nk_priv.p;
            p.W = W;
            HIP_TRY(hipEventRecord(c->ev0, s));
            HIP_TRY(lda::launch_sweep_copy_w(c->kpl, p, s));
            HIP_TRY(hipEventRecord(c->ev1, s));
            HIP_TRY(lda::launch_merge_copies_f(reinterpret_cast<float*>(c->d_nkw.p),
                                               reinterpret_cast<const float*>(c->d_nkw_loc.p), n, W, s));
            c->stat_launches += 1;
        } else if (W == 0) {
            HIP_TRY(lda::launch_fill_u32(c->d_counter.p, 0u, 1, s));
            p.nk_in = c->d_nk.p;
            p.nk_out = c->d_nk.p;
            p.doc_order = c->d_order.p;
            p.counter = c->d_counter.p;
            int waves = 1;
            int rows = 0;
            const int grid = device_wide_grid(c, &waves, &rows);
            if (c->kpl > lda::kSmallMaxKpl || p.exact) {
                HIP_TRY(c->d_bk_scratch.alloc((size_t)grid * std::max(2, waves) * c->Kp));
                p.bk_scratch = c->d_bk_scratch.p;
            }
            p.bk_rows = rows;
            p.col_group = device_col_group(c, (int64_t)grid * waves);
            dw = c->dw_next;
            c->dw_next ^= 1;
            HIP_TRY(hipEventRecord(c->dw_ev[dw][0], s));
            HIP_TRY(lda::launch_sweep_device_w(c->kpl, p, grid, waves, s));
            HIP_TRY(hipEventRecord(c->dw_ev[dw][1], s));
            c->stat_launches += 1;
        } else {
            int rc = prepare_partition(c, W);
            if (rc) return rc;
            HIP_TRY(lda::launch_replicate(c->d_nk_priv.p, c->d_nk.p, c->Kp, W, s));
            p.nk_in = c->d_nk_priv.p;
            p.nk_out = c->d_nk_priv.p;
            p.coff = c->d_coff.p;
            p.W = W;
            HIP_TRY(hipEventRecord(c->ev0, s));
            rc = run_partition_rounds(c, p, W, true);
            if (rc) return rc;
            HIP_TRY(hipEventRecord(c->ev1, s));
            c->stat_launches += W;
        }
        HIP_TRY(lda::launch_clamp_rowsum_f(reinterpret_cast<float*>(c->d_nkw.p), c->V, c->Kp, c->d_rowpart.p,
                                           reinterpret_cast<float*>(c->d_nk.p), s));
        if (dw >= 0) {
            // the previous sweep's time; this one's is read after the next is queued
            if (int rc = dw_read(c)) return rc;
            c->dw_pending = dw;
        } else {
            HIP_TRY(hipEventSynchronize(c->ev1));
            HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            c->stat_ms += ms;
            HIP_TRY(hipStreamSynchronize(s));
        }
        c->stat_tokens += (double)c->n_valid;
        return LDA_OK;
    }
    if (copy_merge) {
        const int64_t n = (int64_t)c->V * c->Kp;
        HIP_TRY(c->d_nkw_loc.alloc((size_t)W * std::max<int64_t>(n, 1)));
        HIP_TRY(c->d_nk_priv.alloc((size_t)W * c->Kp));
        c->part_W = 0;  // d_nk_priv now holds copy_merge state
        HIP_TRY(lda::launch_replicate(c->d_nkw_loc.p, c->d_nkw.p, n, W, s));
        HIP_TRY(lda::launch_replicate(c->d_nk_priv.p, c->d_nk.p, c->Kp, W, s));
        p.nkw = c->d_nkw_loc.p;
        p.nk_in = c->d_nk_priv.p;
        p.nk_out = c->d_nk_priv.p;
        p.W = W;
        HIP_TRY(hipEventRecord(c->ev0, s));
        HIP_TRY(lda::launch_sweep_copy(c->kpl, p, s));
        HIP_TRY(hipEventRecord(c->ev1, s));
        HIP_TRY(lda::launch_merge_copies(c->d_nkw.p, c->d_nkw_loc.p, n, W, s));
        HIP_TRY(lda::launch_rowsum_nkw(c->d_nkw.p, c->V, c->Kp, c->d_nk.p, s));
        HIP_TRY(hipEventSynchronize(c->ev1));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->stat_ms += ms;
        c->stat_launches += 1;
        HIP_TRY(hipStreamSynchronize(s));
        c->stat_tokens += (double)c->n_valid;
        return LDA_OK;
    }
    if (W == 0) {
        HIP_TRY(lda::launch_copy_i32(c->d_nk_next.p, c->d_nk.p, c->Kp, s));
        HIP_TRY(lda::launch_fill_u32(c->d_counter.p, 0u, 1, s));
        p.nk_in = c->d_nk.p;
        p.nk_out = c->d_nk_next.p;
        p.doc_order = c->d_order.p;
        p.counter = c->d_counter.p;
        int waves = 1;
        int rows = 0;
        const int grid = device_wide_grid(c, &waves, &rows);
        if (c->kpl > lda::kSmallMaxKpl || p.exact) {
            HIP_TRY(c->d_bk_scratch.alloc((size_t)grid * std::max(2, waves) * c->Kp));
            p.bk_scratch = c->d_bk_scratch.p;
        }
        p.bk_rows = rows;
        p.col_group = device_col_group(c, (int64_t)grid * waves);
        const int dw = c->dw_next;
        c->dw_next ^= 1;
        HIP_TRY(hipEventRecord(c->dw_ev[dw][0], s));
        HIP_TRY(lda::launch_sweep_device(c->kpl, p, grid, waves, s));
        HIP_TRY(hipEventRecord(c->dw_ev[dw][1], s));
        // the previous sweep's time; this one's is read after the next is queued
        if (int rc = dw_read(c)) return rc;
        c->dw_pending = dw;
        c->stat_launches += 1;
        std::swap(c->d_nk, c->d_nk_next);
    } else {
        int rc = prepare_partition(c, W);
        if (rc) return rc;
        HIP_TRY(lda::launch_replicate(c->d_nk_priv.p, c->d_nk.p, c->Kp, W, s));
        p.nk_in = c->d_nk_priv.p;
        p.nk_out = c->d_nk_priv.p;
        p.coff = c->d_coff.p;
        p.W = W;
        HIP_TRY(hipEventRecord(c->ev0, s));
        rc = run_partition_rounds(c, p, W, false);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(c->ev1, s));
        // mergeState<partition>: n_k = rowwise sum of n_kw (LDAModel.hpp:571-579)
        HIP_TRY(lda::launch_rowsum_nkw(c->d_nkw.p, c->V, c->Kp, c->d_nk.p, s));
        HIP_TRY(hipEventSynchronize(c->ev1));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->stat_ms += ms;
        c->stat_launches += W;
        HIP_TRY(hipStreamSynchronize(s));
    }
    c->stat_tokens += (double)c->n_valid;
    return LDA_OK;
}

// optimizeParameters (LDAModel.hpp:221-265): 10 fixed-point passes; per pass
// the device sums digammaApprox(n_dk + alpha_k) over documents for every
// topic and digammaApprox(L_d + sum alpha) for the denominator (double
// accumulation of float terms); dAlpha uses tomoto's LUT digamma.
int optimize_alpha(lda_ctx* c) {
    if (!c->D) return LDA_OK;
    const hipStream_t s = c->stream;
    const DigammaLut& dg = digamma_lut();
    if (!c->d_dglut.p) {
        HIP_TRY(c->d_dglut.alloc(dg.pts.size()));
        HIP_TRY(hipMemcpyAsync(c->d_dglut.p, dg.pts.data(), dg.pts.size() * sizeof(float), hipMemcpyHostToDevice, s));
    }
    HIP_TRY(c->d_aux.alloc(2));
    HIP_TRY(c->d_bad.alloc(1));
    HIP_TRY(lda::launch_fill_u32(c->d_bad.p, 0u, 1, s));
    float* asum = c->d_aux.p;
    float* den = c->d_aux.p + 1;
    const int Lmax = (int)c->max_len;
    // the histogram form for integer counts while its table stays small
    const bool hist = c->tw == LDA_TW_ONE && (int64_t)c->Kp * (c->max_len + 1) <= (int64_t)1 << 26;
    int nblk = 0;
    if (hist) {
        const size_t nh = (size_t)c->Kp * (Lmax + 1);
        HIP_TRY(c->d_hist.alloc(nh));
        HIP_TRY(c->d_hist_len.alloc(Lmax + 1));
        HIP_TRY(lda::launch_fill_u32(reinterpret_cast<uint32_t*>(c->d_hist.p), 0u, (int64_t)nh, s));
        HIP_TRY(lda::launch_fill_u32(reinterpret_cast<uint32_t*>(c->d_hist_len.p), 0u, Lmax + 1, s));
        HIP_TRY(lda::launch_alpha_hist(c->d_ndk.p, c->d_doclen.p, c->D, c->Kp, Lmax, c->d_hist.p, c->d_hist_len.p, s));
    } else {
        nblk = (int)std::min<int64_t>(std::max<int64_t>(c->D / 32, 1), 2048);
        HIP_TRY(c->d_part.alloc((size_t)nblk * (c->Kp + 1)));
    }
    for (int it = 0; it < 10; ++it) {
        HIP_TRY(lda::launch_alpha_asum(c->d_alpha.p, c->K, c->kpl, asum, s));
        if (hist) {
            HIP_TRY(lda::launch_alpha_den_hist(c->d_hist_len.p, Lmax, c->D, asum, c->d_dglut.p, den, s));
            HIP_TRY(lda::launch_alpha_update_hist(c->d_alpha.p, c->d_hist.p, Lmax, c->Kp, c->D, den, c->d_dglut.p,
                                                  kMaxAlpha, c->d_bad.p, s));
        } else {
            if (c->tw != LDA_TW_ONE)
                HIP_TRY(lda::launch_alpha_sums_f(reinterpret_cast<const float*>(c->d_ndk.p), c->d_docw.p, c->D,
                                                 c->d_alpha.p, c->Kp, asum, c->d_part.p, nblk, s));
            else
                HIP_TRY(lda::launch_alpha_sums(c->d_ndk.p, c->d_doclen.p, c->D, c->d_alpha.p, c->Kp, asum, c->d_part.p,
                                               nblk, s));
            HIP_TRY(lda::launch_alpha_update_direct(c->d_alpha.p, c->d_part.p, nblk, c->Kp, c->D, asum, c->d_dglut.p,
                                                    kMaxAlpha, c->d_bad.p, s));
        }
    }
    std::vector<float> a(c->Kp);
    uint32_t bad = 0;
    HIP_TRY(hipMemcpyAsync(a.data(), c->d_alpha.p, c->Kp * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&bad, c->d_bad.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int k = 0; k < c->K; ++k) c->alpha[k] = a[c->pos(k)];
    // an alpha reached 2^20: integer counts sample on the exact-count kernels from now on
    if (bad) c->exact_alpha = true;
    return upload_alpha(c);  // the cumulative sums for the initial draws
}

// n_kw (V x K, word-major) and n_k into the lane layout; T is int32_t or
// float, the count type of the model.
template <typename T>
int set_counts_t(lda_ctx* c, const T* nkw, const T* nk) {
    HIP_TRY(hipSetDevice(c->device));
    std::vector<T> a((size_t)c->V * c->Kp, T(0)), b(c->Kp, T(0));
    for (int v = 0; v < c->V; ++v)
        for (int k = 0; k < c->K; ++k) a[(size_t)v * c->Kp + c->pos(k)] = nkw[(size_t)v * c->K + k];
    for (int k = 0; k < c->K; ++k) b[c->pos(k)] = nk[k];
    if (!a.empty()) HIP_TRY(hipMemcpyAsync(c->d_nkw.p, a.data(), a.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_nk.p, b.data(), b.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LDA_OK;
}

}  // namespace

namespace {

// Device buffers of one batched / plain fold-in call: the token arena
// (words, z), the jobs, the parameter block, the bucket counters and the
// outputs.
struct FoldArena {
    uint32_t* words;
    uint16_t* z;
    lda::InferJob* jobs;
    lda::InferParams* params;
    uint32_t* counters;
    float* theta;
    double* ll;
};

// Jobs of a fold-in call grouped by bucket (segments per document), longest
// first inside a bucket (the persistent kernels take them in this order), and
// the parameter block pointing into the arena.  mode 1 keeps document order.
void build_fold_jobs(lda_ctx* c, const uint64_t* doc_off, int64_t D, int iterations, uint64_t infer_seed, int mode,
                     const FoldArena& a, std::vector<int>& segs, std::vector<lda::InferJob>& jobs,
                     lda::InferParams& p) {
    segs.resize(D);
    for (int64_t d = 0; d < D; ++d) segs[d] = lda::fold_segments(doc_off[d + 1] - doc_off[d], c->seg_tokens, c->seg_max);
    std::vector<uint32_t> order(D);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
        if (segs[x] != segs[y]) return segs[x] < segs[y];
        return (doc_off[x + 1] - doc_off[x]) > (doc_off[y + 1] - doc_off[y]);
    });
    jobs.resize(D);
    p = lda::InferParams{};
    for (int64_t i = 0; i < D; ++i) {
        const uint32_t d = mode == 0 ? order[i] : (uint32_t)i;
        jobs[i] = lda::InferJob{doc_off[d], doc_off[d + 1], 0, iterations, iterations, (int32_t)d};
    }
    for (int bk = 0; bk < lda::kFoldBuckets; ++bk) p.bucket_lo[bk] = p.bucket_hi[bk] = 0;
    if (mode == 0) {
        int64_t i = 0;
        for (int bk = 0; bk < lda::kFoldBuckets; ++bk) {
            p.bucket_lo[bk] = i;
            while (i < D && lda::fold_bucket(segs[order[i]]) == bk) ++i;
            p.bucket_hi[bk] = i;
        }
    }
    p.words = a.words; p.z = a.z; p.jobs = a.jobs; p.njobs = D;
    p.nkw = c->d_nkw.p; p.nk = c->d_nk.p; p.alpha = c->d_alpha.p; p.alpha_cdf = c->d_alpha_cdf.p;
    p.eta = c->eta; p.veta = c->veta; p.K = c->K; p.V = c->V; p.Kp = c->Kp;
    p.seed_lo = (uint32_t)infer_seed; p.seed_hi = (uint32_t)(infer_seed >> 32);
    p.theta = a.theta; p.ll = a.ll;
    double asd = 0.0;
    float asf = 0.0f;
    for (int k = 0; k < c->K; ++k) { asd += (double)c->alpha[k]; asf = asf + c->alpha[k]; }
    p.asum_d = asd; p.asum_f = asf;
    if (c->has_priors) {
        p.prior_row = c->d_prior_row.p; p.prior_eta = c->d_prior_eta.p;
        p.etasum = c->d_etasum.p; p.etasum_d = c->d_etasum_d.p;
    }
    p.counters = mode == 0 ? a.counters : nullptr;
    p.job_base = 0;
}

// Outputs of a finished fold-in call back to the host.
int fold_download(lda_ctx* c, const FoldArena& a, int64_t D, uint64_t N, int iterations, float* theta, double* ll,
                  uint16_t* z_out) {
    const hipStream_t s = c->stream;
    HIP_TRY(hipEventSynchronize(c->ev1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stat_ms = ms;
    c->stat_tokens = (double)N * iterations;
    HIP_TRY(hipMemcpyAsync(theta, a.theta, (size_t)D * c->K * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(ll, a.ll, D * 8, hipMemcpyDeviceToHost, s));
    if (z_out && N) HIP_TRY(hipMemcpyAsync(z_out, a.z, N * 2, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return LDA_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Fold-in without allocation (lda_infer_noalloc).  Every device buffer of the
// call lives in `scratch`, a file-scope device array of kNoAllocScratchBytes
// that each GPU holds once for the life of the process.  Host code never uses
// the array's name as an address: its device address is taken once per
// device with hipGetSymbolAddress (at context creation, with that device
// current) and kept in g_scratch_dev.
namespace {
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

int noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(LDA_ERR_DEVICE, "device %d beyond the scratch table", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[device]) return LDA_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(LDA_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch_dev[device] = ptr;
    return LDA_OK;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

namespace {
// Device-to-device copy of a whole buffer (its allocated capacity).
template <typename T>
hipError_t dup_buf(DevBuf<T>& dst, const DevBuf<T>& src, hipStream_t s) {
    if (!src.p) return hipSuccess;
    hipError_t e = dst.alloc(src.n);
    if (e == hipSuccess) e = hipMemcpyAsync(dst.p, src.p, src.n * sizeof(T), hipMemcpyDeviceToDevice, s);
    return e;
}
}  // namespace

// ============================================================== C ABI ======
extern "C" {

const char* lda_last_error(void) { return lda_host::g_err.c_str(); }

int lda_version(void) { return 1; }

int lda_device_count(int* count) {
    if (!count) return fail(LDA_ERR_INVALID_ARGUMENT, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = (e == hipSuccess) ? n : 0;
    return LDA_OK;
}

int lda_ctx_create(int K, int V, const float* alpha, float eta, int termweight, uint64_t seed, int device,
                   lda_ctx** out) {
    if (!out) return fail(LDA_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    // LDAModel ctor checks (LDAModel.hpp:964-984)
    if (K <= 0 || K > LDA_MAX_K) return fail(LDA_ERR_INVALID_ARGUMENT, "wrong K value (K = %d); the device sampler supports 1..%d", K, LDA_MAX_K);
    if (V < 0) return fail(LDA_ERR_INVALID_ARGUMENT, "wrong V value (V = %d)", V);
    if (!alpha) return fail(LDA_ERR_INVALID_ARGUMENT, "alpha is NULL");
    for (int k = 0; k < K; ++k) {
        if (!(alpha[k] > 0.0f)) return fail(LDA_ERR_INVALID_ARGUMENT, "wrong alpha value");
        if (!(alpha[k] < 3.0e38f)) return fail(LDA_ERR_INVALID_ARGUMENT, "wrong alpha value");
    }
    if (!(eta > 0.0f)) return fail(LDA_ERR_INVALID_ARGUMENT, "wrong eta value (eta = %f)", (double)eta);
    if (termweight != LDA_TW_ONE && termweight != LDA_TW_IDF && termweight != LDA_TW_PMI)
        return fail(LDA_ERR_INVALID_ARGUMENT, "invalid TermWeight %d", termweight);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(LDA_ERR_DEVICE, "no HIP device is available");
    if (device < 0 || device >= ndev) return fail(LDA_ERR_INVALID_ARGUMENT, "device %d out of range (%d devices)", device, ndev);
    HIP_TRY(hipSetDevice(device));
    lda_ctx* c = new lda_ctx();
    c->device = device;
    c->K = K;
    c->V = V;
    c->kpl = kpl_for(K);
    c->Kp = 64 * c->kpl;
    c->eta = eta;
    c->veta = eta * (float)V;  // EtaHelper::getEtaSum (LDAModel.hpp:93-96)
    c->tw = termweight;
    c->seed = seed;
    c->alpha.assign(alpha, alpha + K);
    for (int k = 0; k < K; ++k) c->exact_alpha = c->exact_alpha || !(alpha[k] < kMaxAlpha);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cu = prop.multiProcessorCount;
    auto cleanup = [&](int rc) {
        lda_ctx_destroy(c);
        return rc;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return cleanup(fail(LDA_ERR_DEVICE, "hipStreamCreate failed"));
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
        return cleanup(fail(LDA_ERR_DEVICE, "hipEventCreate failed"));
    for (auto& pair : c->dw_ev)
        for (auto& ev : pair)
            if (hipEventCreate(&ev) != hipSuccess) return cleanup(fail(LDA_ERR_DEVICE, "hipEventCreate failed"));
    if (c->fold_streams.create() != hipSuccess) return cleanup(fail(LDA_ERR_DEVICE, "fold-in stream creation failed"));
    compute_fold_grid(c);
    if (int src = noalloc_scratch_init(device)) return cleanup(src);
    if (c->d_alpha.alloc(c->Kp) || c->d_alpha_cdf.alloc(K) || c->d_nkw.alloc((size_t)std::max(V, 1) * c->Kp) ||
        c->d_nk.alloc(c->Kp) || c->d_nk_next.alloc(c->Kp) || c->d_counter.alloc(1))
        return cleanup(fail(LDA_ERR_OUT_OF_MEMORY, "device allocation failed"));
    if (hipMemsetAsync(c->d_nkw.p, 0, c->d_nkw.n * sizeof(int32_t), c->stream) != hipSuccess ||
        hipMemsetAsync(c->d_nk.p, 0, c->Kp * sizeof(int32_t), c->stream) != hipSuccess)
        return cleanup(fail(LDA_ERR_DEVICE, "hipMemsetAsync failed"));
    int rc = upload_alpha(c);
    if (rc) return cleanup(rc);
    *out = c;
    return LDA_OK;
}

int lda_ctx_destroy(lda_ctx* c) {
    if (!c) return LDA_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->d_alpha.release(); c->d_alpha_cdf.release(); c->d_nkw.release(); c->d_nk.release(); c->d_nk_next.release();
    c->d_ndk.release(); c->d_doclen.release(); c->d_nk_priv.release(); c->d_nkw_loc.release(); c->d_words.release(); c->d_order.release();
    c->d_bk_scratch.release();
    c->d_counter.release(); c->d_z.release(); c->d_off.release(); c->d_coff.release(); c->d_part.release();
    if (c->infer_exec) (void)hipGraphExecDestroy(c->infer_exec);
    if (c->part_exec) (void)hipGraphExecDestroy(c->part_exec);
    c->d_sweep.release();
    c->i_words.release(); c->i_counter.release(); c->i_jobs.release(); c->i_params.release(); c->i_z.release();
    c->i_theta.release(); c->i_ll.release();
    c->d_wt.release(); c->d_docw.release(); c->d_rowpart.release();
    c->d_prior_row.release(); c->d_prior_eta.release(); c->d_etasum.release(); c->d_etasum_d.release();
    c->gen.release();
    c->d_dglut.release(); c->d_aux.release(); c->d_bad.release(); c->d_hist.release(); c->d_hist_len.release();
    c->d_nkw16.release(); c->r_out.release(); c->r_cnt.release();
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (auto& pair : c->dw_ev)
        for (auto& ev : pair)
            if (ev) (void)hipEventDestroy(ev);
    c->fold_streams.destroy();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return LDA_OK;
}

int lda_ctx_copy(lda_ctx* c, lda_ctx** out) {
    if (!c || !out) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    lda_ctx* t = nullptr;
    int rc = lda_ctx_create(c->K, c->V, c->alpha.data(), c->eta, c->tw, c->seed, c->device, &t);
    if (rc) return rc;
    auto cleanup = [&](int code) {
        lda_ctx_destroy(t);
        return code;
    };
    if (hipStreamSynchronize(c->stream) != hipSuccess) return cleanup(fail(LDA_ERR_DEVICE, "hipStreamSynchronize failed"));
    t->global_step = c->global_step;
    t->optim_interval = c->optim_interval;
    t->burn_in = c->burn_in;
    t->D = c->D;
    t->N = c->N;
    t->n_valid = c->n_valid;
    t->off = c->off;
    t->words = c->words;
    t->has_corpus = c->has_corpus;
    t->wt = c->wt;
    t->weighted_n = c->weighted_n;
    t->has_priors = c->has_priors;
    t->chunk_cf = c->chunk_cf;
    t->max_len = c->max_len;
    t->seg_tokens = c->seg_tokens;
    t->seg_max = c->seg_max;
    t->narrow_enabled = c->narrow_enabled;
    t->nkw_hot = c->nkw_hot;
    t->dev_auto = c->dev_auto;
    t->dev_workers = c->dev_workers;
    t->dev_waves = c->dev_waves;
    t->dev_col_group = c->dev_col_group;
    const hipStream_t s = t->stream;
    hipError_t e = hipSuccess;
    // the count buffers first (lda_ctx_create sized them; copy exactly)
    if (e == hipSuccess) e = hipMemcpyAsync(t->d_nkw.p, c->d_nkw.p, c->d_nkw.n * sizeof(int32_t), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(t->d_nk.p, c->d_nk.p, c->Kp * sizeof(int32_t), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = dup_buf(t->d_words, c->d_words, s);
    if (e == hipSuccess) e = dup_buf(t->d_z, c->d_z, s);
    if (e == hipSuccess) e = dup_buf(t->d_off, c->d_off, s);
    if (e == hipSuccess) e = dup_buf(t->d_ndk, c->d_ndk, s);
    if (e == hipSuccess) e = dup_buf(t->d_doclen, c->d_doclen, s);
    if (e == hipSuccess) e = dup_buf(t->d_order, c->d_order, s);
    if (e == hipSuccess) e = dup_buf(t->d_wt, c->d_wt, s);
    if (e == hipSuccess) e = dup_buf(t->d_docw, c->d_docw, s);
    if (e == hipSuccess && c->d_rowpart.p) e = t->d_rowpart.alloc(c->d_rowpart.n);
    if (e == hipSuccess) e = dup_buf(t->d_prior_row, c->d_prior_row, s);
    if (e == hipSuccess) e = dup_buf(t->d_prior_eta, c->d_prior_eta, s);
    if (e == hipSuccess) e = dup_buf(t->d_etasum, c->d_etasum, s);
    if (e == hipSuccess) e = dup_buf(t->d_etasum_d, c->d_etasum_d, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return cleanup(fail(e == hipErrorOutOfMemory ? LDA_ERR_OUT_OF_MEMORY : LDA_ERR_DEVICE, "model copy: %s",
                            hipGetErrorString(e)));
    compute_fold_grid(t);
    *out = t;
    return LDA_OK;
}

int lda_debug_fail_at_step(lda_ctx* c, int64_t step) {
    if (!c) return fail(LDA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    c->fail_at_step = step < 0 ? -1 : step;
    return LDA_OK;
}

int lda_upload_corpus(lda_ctx* c, const uint32_t* words, const uint64_t* doc_off, int64_t D, const uint16_t* z,
                      const float* weights) {
    if (!c) return fail(LDA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    const bool weighted = c->tw != LDA_TW_ONE;
    if (weighted && D && !weights)
        return fail(LDA_ERR_INVALID_ARGUMENT, "a term-weighted model needs token weights");
    if (D < 0 || !doc_off) return fail(LDA_ERR_INVALID_ARGUMENT, "invalid document offsets");
    if (D > 0xFFFFFFFFll) return fail(LDA_ERR_INVALID_ARGUMENT, "too many documents");
    if (doc_off[0] != 0) return fail(LDA_ERR_INVALID_ARGUMENT, "doc_off[0] must be 0");
    for (int64_t d = 0; d < D; ++d) {
        if (doc_off[d + 1] < doc_off[d]) return fail(LDA_ERR_INVALID_ARGUMENT, "doc_off must be non-decreasing");
    }
    const uint64_t N = doc_off[D];
    if (N && !words) return fail(LDA_ERR_INVALID_ARGUMENT, "words is NULL");
    for (int64_t d = 0; d < D; ++d)
        for (uint64_t g = doc_off[d] + 1; g < doc_off[d + 1]; ++g)
            if (words[g] < words[g - 1]) return fail(LDA_ERR_INVALID_ARGUMENT, "words must be sorted within each document");
    uint64_t nvalid = 0;
    int64_t max_len = 0;
    for (int64_t d = 0; d < D; ++d) {
        int64_t L = 0;
        for (uint64_t g = doc_off[d]; g < doc_off[d + 1]; ++g) L += words[g] < (uint32_t)c->V;
        max_len = std::max(max_len, L);
    }
    for (uint64_t g = 0; g < N; ++g) {
        if (words[g] < (uint32_t)c->V) {
            ++nvalid;
            if (z && z[g] >= (uint16_t)c->K) return fail(LDA_ERR_INVALID_ARGUMENT, "z out of range at token %llu", (unsigned long long)g);
            if (weighted && !(weights[g] >= 0.0f && weights[g] < 1e30f))
                return fail(LDA_ERR_INVALID_ARGUMENT, "token weight %llu is negative or not finite", (unsigned long long)g);
        }
    }
    HIP_TRY(hipSetDevice(c->device));
    c->D = D;
    c->N = N;
    c->n_valid = nvalid;
    c->max_len = max_len;
    c->off.assign(doc_off, doc_off + D + 1);
    c->words.assign(words, words + N);
    c->part_W = 0;
    const hipStream_t s = c->stream;
    HIP_TRY(c->d_words.alloc(N));
    HIP_TRY(c->d_z.alloc(N));
    HIP_TRY(c->d_off.alloc(D + 1));
    HIP_TRY(c->d_ndk.alloc((size_t)std::max<int64_t>(D, 1) * c->Kp));
    HIP_TRY(c->d_doclen.alloc(std::max<int64_t>(D, 1)));
    HIP_TRY(c->d_order.alloc(std::max<int64_t>(D, 1)));
    if (N) HIP_TRY(hipMemcpyAsync(c->d_words.p, words, N * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->d_off.p, doc_off, (D + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    if (z) {
        if (N) HIP_TRY(hipMemcpyAsync(c->d_z.p, z, N * sizeof(uint16_t), hipMemcpyHostToDevice, s));
    } else {
        // initial topics (LDAModel.hpp:754-762) from the Philox stream
        // (with word priors: from the word's eta column, LDAModel.hpp:769-773)
        HIP_TRY(lda::launch_init_topics(c->d_words.p, c->d_z.p, N, c->d_alpha_cdf.p, c->K, c->V, (uint32_t)c->seed,
                                        (uint32_t)(c->seed >> 32), c->has_priors ? c->d_prior_row.p : nullptr,
                                        c->d_prior_eta.p, c->kpl, c->Kp, c->eta, s));
    }
    if (!weighted) {
        HIP_TRY(hipMemsetAsync(c->d_ndk.p, 0, c->d_ndk.n * sizeof(int32_t), s));
        HIP_TRY(hipMemsetAsync(c->d_nkw.p, 0, c->d_nkw.n * sizeof(int32_t), s));
        HIP_TRY(hipMemsetAsync(c->d_nk.p, 0, c->Kp * sizeof(int32_t), s));
        HIP_TRY(lda::launch_build_counts(c->kpl, c->d_words.p, c->d_z.p, c->d_off.p, D, c->d_ndk.p, c->d_nkw.p, c->d_nk.p,
                                         c->Kp, c->V, s));
        HIP_TRY(lda::launch_doclen(c->d_words.p, c->d_off.p, D, c->V, c->d_doclen.p, s));
    } else {
        // float counts: addWordTo<+w> in token order (updateStateWithDoc,
        // LDAModel.hpp:776-813), accumulated on the host so that the sums
        // follow tomoto's order; per-document weight sums (getSumWordWeight)
        c->wt.assign(weights, weights + N);
        std::vector<uint16_t> zh(N);
        if (z) std::copy(z, z + N, zh.begin());
        else if (N) {
            HIP_TRY(hipMemcpyAsync(zh.data(), c->d_z.p, N * sizeof(uint16_t), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
        }
        std::vector<float> ndk((size_t)std::max<int64_t>(D, 1) * c->Kp, 0.0f), nkw((size_t)c->V * c->Kp, 0.0f),
            nk(c->Kp, 0.0f), docw(std::max<int64_t>(D, 1), 0.0f);
        if (c->base_nkw.size() == nkw.size() && c->base_nk.size() == nk.size()) {
            nkw.swap(c->base_nkw);
            nk.swap(c->base_nk);
        }
        c->base_nkw.clear();
        c->base_nk.clear();
        c->weighted_n = 0.0;
        for (int64_t d = 0; d < D; ++d) {
            float sw = 0.0f;
            for (uint64_t g = doc_off[d]; g < doc_off[d + 1]; ++g) {
                const uint32_t v = words[g];
                if (v >= (uint32_t)c->V) continue;
                const float w = weights[g];
                const int q = c->pos(zh[g]);
                ndk[(size_t)d * c->Kp + q] += w;
                nkw[(size_t)v * c->Kp + q] += w;
                nk[q] += w;
                sw += w;
                c->weighted_n += (double)w;  // countRealN (TopicModel.hpp:507-524)
            }
            docw[d] = sw;
        }
        HIP_TRY(c->d_wt.alloc(std::max<uint64_t>(N, 1)));
        HIP_TRY(c->d_rowpart.alloc((size_t)std::max(1, (c->V + lda::kRowsumBlock - 1) / lda::kRowsumBlock) * c->Kp));
        HIP_TRY(c->d_docw.alloc(std::max<int64_t>(D, 1)));
        if (N) HIP_TRY(hipMemcpyAsync(c->d_wt.p, weights, N * sizeof(float), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(c->d_docw.p, docw.data(), docw.size() * sizeof(float), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(c->d_ndk.p, ndk.data(), ndk.size() * sizeof(float), hipMemcpyHostToDevice, s));
        if (!nkw.empty()) HIP_TRY(hipMemcpyAsync(c->d_nkw.p, nkw.data(), nkw.size() * sizeof(float), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(c->d_nk.p, nk.data(), nk.size() * sizeof(float), hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    // the most frequent words keep int32 columns in the narrow-column sweep:
    // every word up to the last one with kNarrowCf or more tokens
    {
        std::vector<uint64_t> cf(c->V, 0);
        for (uint64_t g = 0; g < N; ++g)
            if (words[g] < (uint32_t)c->V) ++cf[words[g]];
        int hot = 0;
        for (int v = 0; v < c->V; ++v)
            if (cf[v] >= lda::kNarrowCf) hot = v + 1;
        c->nkw_hot = hot;
    }
    // documents longest first for the device-wide schedule
    std::vector<uint32_t> order(D);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return (doc_off[a + 1] - doc_off[a]) > (doc_off[b + 1] - doc_off[b]);
    });
    if (D) HIP_TRY(hipMemcpyAsync(c->d_order.p, order.data(), D * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    c->has_corpus = true;
    return LDA_OK;
}


int lda_set_counts(lda_ctx* c, const int32_t* nkw, const int32_t* nk) {
    if (c && c->tw != LDA_TW_ONE) return fail(LDA_ERR_UNSUPPORTED, "lda_set_counts takes integer counts (TermWeight ONE)");
    if (!c || !nkw || !nk) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    return set_counts_t<int32_t>(c, nkw, nk);
}

int lda_set_counts_f(lda_ctx* c, const float* nkw, const float* nk) {
    if (c && c->tw == LDA_TW_ONE) return fail(LDA_ERR_STATE, "integer counts: use lda_set_counts");
    if (!c || !nkw || !nk) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    for (size_t i = 0; i < (size_t)c->V * c->K; ++i)
        if (!(nkw[i] >= 0.0f && nkw[i] < 1e30f)) return fail(LDA_ERR_INVALID_ARGUMENT, "n_kw must be finite and >= 0");
    return set_counts_t<float>(c, nkw, nk);
}

int lda_set_word_priors(lda_ctx* c, int64_t n, const uint32_t* word_ids, const float* priors) {
    if (!c || n < 0 || (n && (!word_ids || !priors))) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    if (c->has_corpus) return fail(LDA_ERR_STATE, "word priors must be set before the corpus is uploaded");
    const int K = c->K, Kp = c->Kp, V = c->V;
    std::vector<int32_t> row(std::max(V, 1), -1);
    std::vector<float> tab;
    int32_t rows = 0;
    for (int64_t i = 0; i < n; ++i) {
        const uint32_t v = word_ids[i];
        for (int k = 0; k < K; ++k) {
            const float x = priors[(size_t)i * K + k];
            if (!(x >= 0.0f && x < 1e30f)) return fail(LDA_ERR_INVALID_ARGUMENT, "priors must not be less than 0.");
        }
        if (v >= (uint32_t)V) continue;  // removed or unknown words are skipped (LDAModel.hpp:733)
        if (row[v] >= 0) return fail(LDA_ERR_INVALID_ARGUMENT, "word %u has two priors", v);
        row[v] = rows++;
        tab.resize((size_t)rows * Kp, 0.0f);
        for (int k = 0; k < K; ++k) tab[(size_t)(rows - 1) * Kp + c->pos(k)] = priors[(size_t)i * K + k];
    }
    if (n == 0) {
        c->has_priors = false;
        compute_fold_grid(c);
        return LDA_OK;
    }
    // etaSumByTopic = rowwise sum of eta_kv (LDAModel.hpp:736): float, words in
    // ascending order (the sampler's), and double (the FP64 log-likelihood's)
    std::vector<float> es(Kp, 1.0f);
    std::vector<double> esd(Kp, 1.0);
    for (int k = 0; k < K; ++k) {
        const int q = c->pos(k);
        float sf = 0.0f;
        double sd = 0.0;
        for (int v = 0; v < V; ++v) {
            const float x = row[v] >= 0 ? tab[(size_t)row[v] * Kp + q] : c->eta;
            sf += x;
            sd += (double)x;
        }
        es[q] = sf;
        esd[q] = sd;
    }
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(c->d_prior_row.alloc(row.size()));
    HIP_TRY(c->d_prior_eta.alloc(std::max<size_t>(tab.size(), 1)));
    HIP_TRY(c->d_etasum.alloc(Kp));
    HIP_TRY(c->d_etasum_d.alloc(Kp));
    HIP_TRY(hipMemcpyAsync(c->d_prior_row.p, row.data(), row.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    if (!tab.empty())
        HIP_TRY(hipMemcpyAsync(c->d_prior_eta.p, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_etasum.p, es.data(), Kp * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_etasum_d.p, esd.data(), Kp * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->has_priors = true;
    compute_fold_grid(c);
    return LDA_OK;
}

int lda_set_alpha(lda_ctx* c, const float* alpha) {
    if (!c || !alpha) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    for (int k = 0; k < c->K; ++k) {
        if (!(alpha[k] > 0.0f)) return fail(LDA_ERR_INVALID_ARGUMENT, "wrong alpha value");
        if (!(alpha[k] < 3.0e38f)) return fail(LDA_ERR_INVALID_ARGUMENT, "wrong alpha value");
    }
    c->alpha.assign(alpha, alpha + c->K);
    c->exact_alpha = false;
    for (int k = 0; k < c->K; ++k) c->exact_alpha = c->exact_alpha || !(alpha[k] < kMaxAlpha);
    HIP_TRY(hipSetDevice(c->device));
    return upload_alpha(c);
}

int lda_get_alpha(lda_ctx* c, float* alpha) {
    if (!c || !alpha) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    std::copy(c->alpha.begin(), c->alpha.end(), alpha);
    return LDA_OK;
}

int lda_set_optim(lda_ctx* c, uint32_t optim_interval, uint32_t burn_in) {
    if (!c) return fail(LDA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (optim_interval > 0x7FFFFFFFu || burn_in > 0x7FFFFFFFu) return fail(LDA_ERR_INVALID_ARGUMENT, "wrong value");
    c->optim_interval = optim_interval;
    c->burn_in = burn_in;
    return LDA_OK;
}

int lda_get_global_step(lda_ctx* c, uint64_t* step) {
    if (!c || !step) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    *step = c->global_step;
    return LDA_OK;
}

int lda_set_global_step(lda_ctx* c, uint64_t step) {
    if (!c) return fail(LDA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    c->global_step = step;
    return LDA_OK;
}

int lda_set_device_schedule(lda_ctx* c, int max_workers, int waves_per_worker) {
    if (!c) return fail(LDA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (max_workers < 0) {
        c->dev_auto = true;
        return LDA_OK;
    }
    if (waves_per_worker < 1 || waves_per_worker > lda::kMaxSweepWaves)
        return fail(LDA_ERR_INVALID_ARGUMENT, "waves_per_worker must be in [1, %d]", lda::kMaxSweepWaves);
    c->dev_auto = false;
    c->dev_workers = max_workers;
    c->dev_waves = waves_per_worker;
    return LDA_OK;
}

int lda_resolve_schedule(lda_ctx* c, int workers, int scheme, int* scheme_out, int* workers_out) {
    if (!c || !scheme_out || !workers_out) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!c->has_corpus) return fail(LDA_ERR_STATE, "no corpus uploaded");
    int W;
    bool cm;
    int rc = resolve_workers(c, workers, scheme, &W, &cm);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(c->device));
    if (W == 0) {
        int waves = 1;
        *scheme_out = LDA_PS_DEFAULT;
        *workers_out = device_wide_grid(c, &waves);
    } else {
        *scheme_out = W == 1 ? LDA_PS_NONE : cm ? LDA_PS_COPY_MERGE : LDA_PS_PARTITION;
        *workers_out = W;
    }
    return LDA_OK;
}

int lda_sweep(lda_ctx* c, int workers, int scheme) {
    if (!c) return fail(LDA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!c->has_corpus) return fail(LDA_ERR_STATE, "no corpus uploaded");
    int W;
    bool cm;
    int rc = resolve_workers(c, workers, scheme, &W, &cm);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(c->device));
    c->stat_ms = c->stat_launches = c->stat_tokens = 0.0;
    rc = narrow_begin(c, W, cm);
    if (rc) return rc;
    rc = sweep_once(c, W, cm);
    const int rc1 = dw_read(c);
    const int rc2 = narrow_end(c);
    return rc ? rc : (rc1 ? rc1 : rc2);
}

int lda_train(lda_ctx* c, int iterations, int workers, int scheme) {
    if (!c) return fail(LDA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (iterations < 0) return fail(LDA_ERR_INVALID_ARGUMENT, "iterations must be >= 0");
    if (!c->has_corpus) return fail(LDA_ERR_STATE, "no corpus uploaded");
    int W;
    bool cm;
    int rc = resolve_workers(c, workers, scheme, &W, &cm);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(c->device));
    c->stat_ms = c->stat_launches = c->stat_tokens = 0.0;
    rc = narrow_begin(c, W, cm);
    if (rc) return rc;
    for (int it = 0; it < iterations && rc == LDA_OK; ++it) {
        if (c->fail_at_step >= 0 && (uint64_t)c->fail_at_step == c->global_step) {
            rc = fail(LDA_ERR_TRAINING, "training error forced at step %llu", (unsigned long long)c->global_step);
            break;
        }
        rc = sweep_once(c, W, cm);
        if (rc) break;
        // trainOne (LDAModel.hpp:529-532)
        if (c->global_step >= c->burn_in && c->optim_interval && (c->global_step + 1) % c->optim_interval == 0) {
            rc = optimize_alpha(c);
            if (rc) break;
        }
        ++c->global_step;
    }
    const int rc1 = dw_read(c);
    const int rc2 = narrow_end(c);
    return rc ? rc : (rc1 ? rc1 : rc2);
}

int lda_optimize_alpha(lda_ctx* c) {
    if (!c) return fail(LDA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!c->has_corpus) return fail(LDA_ERR_STATE, "no corpus uploaded");
    HIP_TRY(hipSetDevice(c->device));
    return optimize_alpha(c);
}

int lda_loglik(lda_ctx* c, double* ll_docs, double* ll_rest) {
    if (!c || !ll_docs || !ll_rest) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    HIP_TRY(hipSetDevice(c->device));
    const int nblk = 1024;
    HIP_TRY(c->d_part.alloc(2 * nblk));
    double asum = 0.0;
    for (int k = 0; k < c->K; ++k) asum += (double)c->alpha[k];
    const double eta = (double)c->eta, veta = (double)c->V * eta;
    const int64_t D = c->has_corpus ? c->D : 0;
    if (c->tw != LDA_TW_ONE)
        HIP_TRY(lda::launch_loglik_f(reinterpret_cast<const float*>(c->d_ndk.p), c->d_docw.p, D,
                                     reinterpret_cast<const float*>(c->d_nkw.p), reinterpret_cast<const float*>(c->d_nk.p),
                                     c->d_alpha.p, c->V, c->Kp, asum, eta, veta, c->d_part.p, nblk,
                                     c->has_priors ? c->d_prior_row.p : nullptr, c->d_prior_eta.p, c->d_etasum_d.p,
                                     c->stream));
    else
        HIP_TRY(lda::launch_loglik(c->d_ndk.p, c->d_doclen.p, D, c->d_nkw.p, c->d_nk.p, c->d_alpha.p, c->V, c->Kp, asum,
                                   eta, veta, c->d_part.p, nblk, c->has_priors ? c->d_prior_row.p : nullptr,
                                   c->d_prior_eta.p, c->d_etasum_d.p, c->stream));
    std::vector<double> part(2 * nblk);
    HIP_TRY(hipMemcpyAsync(part.data(), c->d_part.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    double a = 0.0, b = 0.0;
    for (int i = 0; i < nblk; ++i) {
        a += part[i];
        b += part[nblk + i];
    }
    *ll_docs = a;
    *ll_rest = b;
    return LDA_OK;
}

int lda_infer_together_w(lda_ctx* c, const uint32_t* words, const uint64_t* doc_off, int64_t D, const float* weights,
                         int iterations, uint64_t infer_seed, int workers, int scheme, float* theta, double* ll,
                         uint16_t* z_out) {
    if (!c || !ll || D < 0 || (D && (!doc_off || !theta))) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    if (iterations < 0) return fail(LDA_ERR_INVALID_ARGUMENT, "iterations must be >= 0");
    const bool weighted = c->tw != LDA_TW_ONE;
    if (!weighted && weights) return fail(LDA_ERR_INVALID_ARGUMENT, "TermWeight ONE models take no token weights");
    if (weighted && D && doc_off[D] && !weights)
        return fail(LDA_ERR_INVALID_ARGUMENT, "term-weighted models need the token weights");
    // the documents join a copy of the model: a context of their own whose
    // counts are the model's plus the documents', sampled with alpha fixed
    lda_ctx* t = nullptr;
    int rc = lda_ctx_create(c->K, c->V, c->alpha.data(), c->eta, c->tw, infer_seed, c->device, &t);
    if (rc) return rc;
    struct Guard {
        lda_ctx* t;
        ~Guard() { lda_ctx_destroy(t); }
    } guard{t};
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->has_priors) {
        // the copy keeps eta_kv: initial topics from the eta columns (LDAModel.hpp:764-775)
        HIP_TRY(t->d_prior_row.alloc(c->d_prior_row.n));
        HIP_TRY(t->d_prior_eta.alloc(c->d_prior_eta.n));
        HIP_TRY(t->d_etasum.alloc(c->d_etasum.n));
        HIP_TRY(t->d_etasum_d.alloc(c->d_etasum_d.n));
        HIP_TRY(hipMemcpyAsync(t->d_prior_row.p, c->d_prior_row.p, c->d_prior_row.n * 4, hipMemcpyDeviceToDevice, t->stream));
        HIP_TRY(hipMemcpyAsync(t->d_prior_eta.p, c->d_prior_eta.p, c->d_prior_eta.n * 4, hipMemcpyDeviceToDevice, t->stream));
        HIP_TRY(hipMemcpyAsync(t->d_etasum.p, c->d_etasum.p, c->d_etasum.n * 4, hipMemcpyDeviceToDevice, t->stream));
        HIP_TRY(hipMemcpyAsync(t->d_etasum_d.p, c->d_etasum_d.p, c->d_etasum_d.n * 8, hipMemcpyDeviceToDevice, t->stream));
        HIP_TRY(hipStreamSynchronize(t->stream));
        t->has_priors = true;
    }
    t->chunk_cf.assign(c->V, 0);
    for (uint32_t v : c->words)
        if (v < (uint32_t)c->V) ++t->chunk_cf[v];
    if (weighted) {
        // float counts: the documents' tokens are added onto the model's in
        // document and token order (initializeDocState on tmpState)
        t->base_nkw.resize((size_t)c->V * c->Kp);
        t->base_nk.resize(c->Kp);
        if (!t->base_nkw.empty())
            HIP_TRY(hipMemcpy(t->base_nkw.data(), c->d_nkw.p, t->base_nkw.size() * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(t->base_nk.data(), c->d_nk.p, c->Kp * 4, hipMemcpyDeviceToHost));
    }
    std::vector<uint64_t> off0(1, 0);
    rc = lda_upload_corpus(t, words, D ? doc_off : off0.data(), D, nullptr, weighted ? weights : nullptr);
    if (rc) return rc;
    if (!weighted) {
        HIP_TRY(lda::launch_add_i32(t->d_nkw.p, c->d_nkw.p, (int64_t)c->V * c->Kp, t->stream));
        HIP_TRY(lda::launch_add_i32(t->d_nk.p, c->d_nk.p, c->Kp, t->stream));
        HIP_TRY(hipStreamSynchronize(t->stream));
    }
    t->optim_interval = 0;
    rc = lda_train(t, iterations, workers, scheme);
    if (rc) return rc;
    double docs_t = 0, rest_t = 0, docs_c = 0, rest_c = 0;
    rc = lda_loglik(t, &docs_t, &rest_t);
    if (rc) return rc;
    rc = lda_loglik(c, &docs_c, &rest_c);
    if (rc) return rc;
    *ll = rest_t - rest_c + docs_t;
    // theta = (n_dk + alpha) / (weight sum + sum alpha) (LDAModel.hpp:1145-1159)
    float asum = 0.0f;
    for (int k = 0; k < c->K; ++k) asum += c->alpha[k];
    if (weighted) {
        std::vector<float> ndk((size_t)std::max<int64_t>(D, 1) * c->K);
        rc = lda_download_state_f(t, z_out, ndk.data(), nullptr, nullptr);
        if (rc) return rc;
        for (int64_t d = 0; d < D; ++d) {
            float sw = 0.0f;
            for (uint64_t g = doc_off[d]; g < doc_off[d + 1]; ++g)
                if (words[g] < (uint32_t)c->V) sw += weights[g];
            for (int k = 0; k < c->K; ++k)
                theta[(size_t)d * c->K + k] = (ndk[(size_t)d * c->K + k] + c->alpha[k]) / (sw + asum);
        }
        return LDA_OK;
    }
    std::vector<int32_t> ndk((size_t)std::max<int64_t>(D, 1) * c->K);
    rc = lda_download_state(t, z_out, ndk.data(), nullptr, nullptr);
    if (rc) return rc;
    for (int64_t d = 0; d < D; ++d) {
        int32_t L = 0;
        for (int k = 0; k < c->K; ++k) L += ndk[(size_t)d * c->K + k];
        for (int k = 0; k < c->K; ++k)
            theta[(size_t)d * c->K + k] = ((float)ndk[(size_t)d * c->K + k] + c->alpha[k]) / ((float)L + asum);
    }
    return LDA_OK;
}

int lda_infer_together(lda_ctx* c, const uint32_t* words, const uint64_t* doc_off, int64_t D, int iterations,
                       uint64_t infer_seed, int workers, int scheme, float* theta, double* ll, uint16_t* z_out) {
    if (c && c->tw != LDA_TW_ONE)
        return fail(LDA_ERR_INVALID_ARGUMENT, "term-weighted models need the token weights: call lda_infer_together_w");
    return lda_infer_together_w(c, words, doc_off, D, nullptr, iterations, infer_seed, workers, scheme, theta, ll, z_out);
}

int lda_check_invariants(lda_ctx* c, int64_t flags[8]) {
    if (!c || !flags) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!c->has_corpus) return fail(LDA_ERR_STATE, "no corpus uploaded");
    if (c->tw != LDA_TW_ONE)
        return fail(LDA_ERR_UNSUPPORTED, "exact count invariants do not apply to float (term-weighted) counts");
    HIP_TRY(hipSetDevice(c->device));
    const hipStream_t s = c->stream;
    DevBuf<int32_t> sndk, snkw;
    DevBuf<long long> colsum;
    DevBuf<unsigned long long> fl;
    HIP_TRY(sndk.alloc((size_t)std::max<int64_t>(c->D, 1) * c->Kp));
    hipError_t e = snkw.alloc((size_t)std::max(c->V, 1) * c->Kp);
    if (e != hipSuccess) { sndk.release(); HIP_TRY(e); }
    if (colsum.alloc(c->Kp) != hipSuccess || fl.alloc(8) != hipSuccess) {
        sndk.release(); snkw.release(); colsum.release(); fl.release();
        return fail(LDA_ERR_OUT_OF_MEMORY, "scratch allocation failed");
    }
    int rc = LDA_OK;
    std::vector<long long> cs(c->Kp);
    std::vector<int32_t> nk(c->Kp);
    unsigned long long f[8] = {0};
    do {
        if (hipMemsetAsync(sndk.p, 0, sndk.n * sizeof(int32_t), s) || hipMemsetAsync(snkw.p, 0, snkw.n * sizeof(int32_t), s) ||
            hipMemsetAsync(fl.p, 0, 8 * sizeof(unsigned long long), s)) { rc = fail(LDA_ERR_DEVICE, "memset failed"); break; }
        hipError_t le = lda::launch_invariants(c->kpl, c->d_words.p, c->d_z.p, c->d_off.p, c->D, c->d_ndk.p, c->d_nkw.p,
                                               c->d_nk.p, c->d_doclen.p, c->K, c->V, c->Kp, sndk.p, snkw.p, colsum.p, fl.p, s);
        if (le != hipSuccess) { rc = fail(LDA_ERR_DEVICE, "invariant kernels: %s", hipGetErrorString(le)); break; }
        if (hipMemcpyAsync(f, fl.p, sizeof f, hipMemcpyDeviceToHost, s) ||
            hipMemcpyAsync(cs.data(), colsum.p, cs.size() * sizeof(long long), hipMemcpyDeviceToHost, s) ||
            hipMemcpyAsync(nk.data(), c->d_nk.p, nk.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s) ||
            hipStreamSynchronize(s)) { rc = fail(LDA_ERR_DEVICE, "download failed"); break; }
    } while (0);
    sndk.release(); snkw.release(); colsum.release(); fl.release();
    if (rc) return rc;
    long long bad_topics = 0, total = 0;
    for (int i = 0; i < c->Kp; ++i) {
        if (cs[i] != nk[i]) ++bad_topics;
        total += nk[i];
    }
    flags[0] = (int64_t)f[0];
    flags[1] = bad_topics;
    flags[2] = total - (long long)c->n_valid;
    flags[3] = (int64_t)f[3];
    flags[4] = (int64_t)f[4];
    flags[5] = (int64_t)f[5];
    flags[6] = (int64_t)f[6];
    flags[7] = 0;
    return LDA_OK;
}

int lda_download_state_f(lda_ctx* c, uint16_t* z, float* ndk, float* nkw, float* nk) {
    if (!c) return fail(LDA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (c->tw == LDA_TW_ONE) return fail(LDA_ERR_STATE, "integer counts: use lda_download_state");
    HIP_TRY(hipSetDevice(c->device));
    const hipStream_t s = c->stream;
    if (z && c->N) HIP_TRY(hipMemcpyAsync(z, c->d_z.p, c->N * sizeof(uint16_t), hipMemcpyDeviceToHost, s));
    std::vector<float> a, b, d;
    if (ndk && c->D) {
        a.resize((size_t)c->D * c->Kp);
        HIP_TRY(hipMemcpyAsync(a.data(), c->d_ndk.p, a.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    }
    if (nkw && c->V) {
        b.resize((size_t)c->V * c->Kp);
        HIP_TRY(hipMemcpyAsync(b.data(), c->d_nkw.p, b.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    }
    if (nk) {
        d.resize(c->Kp);
        HIP_TRY(hipMemcpyAsync(d.data(), c->d_nk.p, d.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    for (int k = 0; k < c->K; ++k) {
        const int q = c->pos(k);
        if (ndk)
            for (int64_t i2 = 0; i2 < c->D; ++i2) ndk[(size_t)i2 * c->K + k] = a[(size_t)i2 * c->Kp + q];
        if (nkw)
            for (int v = 0; v < c->V; ++v) nkw[(size_t)v * c->K + k] = b[(size_t)v * c->Kp + q];
        if (nk) nk[k] = d[q];
    }
    return LDA_OK;
}

int lda_selftest_reciprocal(int exp_shift, int device, uint64_t* mismatches) {
    if (!mismatches) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    if (exp_shift < -125 || exp_shift > 126) return fail(LDA_ERR_INVALID_ARGUMENT, "exp_shift must give a normal binade");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(LDA_ERR_DEVICE, "no HIP device is available");
    if (device < 0 || device >= ndev) return fail(LDA_ERR_INVALID_ARGUMENT, "device %d out of range", device);
    HIP_TRY(hipSetDevice(device));
    DevBuf<unsigned long long> bad;
    HIP_TRY(bad.alloc(1));
    HIP_TRY(hipMemset(bad.p, 0, sizeof(unsigned long long)));
    HIP_TRY(lda::launch_selftest_reciprocal(exp_shift, bad.p, nullptr));
    unsigned long long h = 0;
    HIP_TRY(hipMemcpy(&h, bad.p, sizeof(h), hipMemcpyDeviceToHost));
    *mismatches = h;
    return LDA_OK;
}

int lda_topic_word_dist(lda_ctx* c, int topic, int normalize, float* out) {
    if (!c || !out) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    if (topic < 0 || topic >= c->K) return fail(LDA_ERR_INVALID_ARGUMENT, "topic_id must < K");
    if (!c->V) return LDA_OK;
    HIP_TRY(hipSetDevice(c->device));
    const hipStream_t s = c->stream;
    const int q = c->pos(topic);
    // sum = n_k + V * eta in float: V * eta as float(V) * eta (LDAModel.hpp:833)
    float sum = 1.0f;
    const float veta = (float)c->V * c->eta;
    if (c->tw != LDA_TW_ONE) {
        float nk = 0.0f;
        HIP_TRY(hipMemcpyAsync(&nk, reinterpret_cast<const float*>(c->d_nk.p) + q, 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        sum = nk + veta;
    } else {
        int32_t nk = 0;
        HIP_TRY(hipMemcpyAsync(&nk, c->d_nk.p + q, 4, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        sum = (float)nk + veta;
    }
    HIP_TRY(c->r_out.alloc(c->V));
    HIP_TRY(lda::launch_topic_word_dist(c->d_nkw.p, c->tw != LDA_TW_ONE, c->V, c->Kp, q, c->eta, sum, normalize,
                                        c->r_out.p, s));
    HIP_TRY(hipMemcpyAsync(out, c->r_out.p, (size_t)c->V * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return LDA_OK;
}

int lda_count_by_topic(lda_ctx* c, uint64_t* out) {
    if (!c || !out) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!c->has_corpus) {
        std::fill(out, out + c->K, (uint64_t)0);
        return LDA_OK;
    }
    HIP_TRY(hipSetDevice(c->device));
    const hipStream_t s = c->stream;
    HIP_TRY(c->r_cnt.alloc(c->K));
    HIP_TRY(lda::launch_fill_u32(reinterpret_cast<uint32_t*>(c->r_cnt.p), 0u, 2 * (int64_t)c->K, s));
    if (c->N)
        HIP_TRY(lda::launch_count_topics(c->d_words.p, c->d_z.p, c->N, c->V, c->K,
                                         reinterpret_cast<unsigned long long*>(c->r_cnt.p), s));
    HIP_TRY(hipMemcpyAsync(out, c->r_cnt.p, (size_t)c->K * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return LDA_OK;
}

int lda_doc_topic_dist(lda_ctx* c, int64_t doc, int normalize, float* out) {
    if (!c || !out) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!c->has_corpus || doc < 0 || doc >= c->D) return fail(LDA_ERR_INVALID_ARGUMENT, "document index out of range");
    HIP_TRY(hipSetDevice(c->device));
    const hipStream_t s = c->stream;
    std::vector<int32_t> row(c->Kp);
    float docw = 0.0f;
    int32_t L = 0;
    HIP_TRY(hipMemcpyAsync(row.data(), c->d_ndk.p + (size_t)doc * c->Kp, (size_t)c->Kp * 4, hipMemcpyDeviceToHost, s));
    if (c->tw != LDA_TW_ONE) HIP_TRY(hipMemcpyAsync(&docw, c->d_docw.p + doc, 4, hipMemcpyDeviceToHost, s));
    else HIP_TRY(hipMemcpyAsync(&L, c->d_doclen.p + doc, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    // _getTopicsByDoc (LDAModel.hpp:1145-1159): (n_dk + alpha) / (weight sum +
    // sum alpha) in float; the weight sum is the valid token count (int32) for
    // TermWeight ONE and the float sum of the token weights otherwise
    float asum = 0.0f;
    for (int k = 0; k < c->K; ++k) asum += c->alpha[k];
    const float den = (c->tw != LDA_TW_ONE ? docw : (float)L) + asum;
    for (int k = 0; k < c->K; ++k) {
        const int32_t bits = row[c->pos(k)];
        float n;
        if (c->tw != LDA_TW_ONE) std::memcpy(&n, &bits, 4);
        else n = (float)bits;
        const float x = n + c->alpha[k];
        out[k] = normalize ? x / den : x;
    }
    return LDA_OK;
}

int lda_probe_bandwidth(int device, uint64_t bytes, int reps, double out[2]) {
    if (!out || reps <= 0 || bytes < (1u << 20) || bytes % 16) return fail(LDA_ERR_INVALID_ARGUMENT, "invalid probe size");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(LDA_ERR_DEVICE, "no HIP device is available");
    if (device < 0 || device >= ndev) return fail(LDA_ERR_INVALID_ARGUMENT, "device %d out of range", device);
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    DevBuf<int32_t> a, b, sink;
    auto done = [&](int rc) {
        a.release(); b.release(); sink.release();
        return rc;
    };
    if (a.alloc(bytes / 4) || b.alloc(bytes / 4) || sink.alloc(1)) return done(fail(LDA_ERR_OUT_OF_MEMORY, "probe buffers"));
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess)
        return done(fail(LDA_ERR_DEVICE, "probe stream / events"));
    auto run1 = [&](bool copy, int grid, double* gbps) -> hipError_t {
        hipError_t e = lda::launch_fill_u32(reinterpret_cast<uint32_t*>(a.p), 1u, (int64_t)(bytes / 4), s);
        for (int w = 0; w < 2 && e == hipSuccess; ++w)
            e = copy ? lda::launch_probe_copy(a.p, b.p, (int64_t)bytes, grid, s)
                     : lda::launch_probe_read(a.p, (int64_t)bytes, sink.p, grid, s);
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        for (int r = 0; r < reps && e == hipSuccess; ++r)
            e = copy ? lda::launch_probe_copy(a.p, b.p, (int64_t)bytes, grid, s)
                     : lda::launch_probe_read(a.p, (int64_t)bytes, sink.p, grid, s);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        *gbps = (copy ? 2.0 : 1.0) * (double)bytes * reps / (ms * 1e-3) / 1e9;
        return e;
    };
    // the best of four grid sizes (4 to 32 workgroups of 256 threads per CU)
    auto run = [&](bool copy, double* gbps) -> hipError_t {
        *gbps = 0.0;
        for (int m = 4; m <= 32; m *= 2) {
            double g = 0.0;
            const hipError_t e = run1(copy, prop.multiProcessorCount * m, &g);
            if (e != hipSuccess) return e;
            *gbps = std::max(*gbps, g);
        }
        return hipSuccess;
    };
    hipError_t e = run(false, &out[0]);
    if (e == hipSuccess) e = run(true, &out[1]);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    if (e != hipSuccess) return done(fail(LDA_ERR_DEVICE, "probe: %s", hipGetErrorString(e)));
    return done(LDA_OK);
}

int lda_get_weighted_n(lda_ctx* c, double* weighted_n) {
    if (!c || !weighted_n) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    *weighted_n = c->tw == LDA_TW_ONE ? (double)c->n_valid : c->weighted_n;
    return LDA_OK;
}

int lda_download_state(lda_ctx* c, uint16_t* z, int32_t* ndk, int32_t* nkw, int32_t* nk) {
    if (!c) return fail(LDA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    HIP_TRY(hipSetDevice(c->device));
    const hipStream_t s = c->stream;
    if (z && c->has_corpus && c->N) {
        HIP_TRY(hipMemcpyAsync(z, c->d_z.p, c->N * sizeof(uint16_t), hipMemcpyDeviceToHost, s));
    }
    if (ndk && c->has_corpus && c->D) {
        std::vector<int32_t> t((size_t)c->D * c->Kp);
        HIP_TRY(hipMemcpyAsync(t.data(), c->d_ndk.p, t.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int64_t d = 0; d < c->D; ++d)
            for (int k = 0; k < c->K; ++k) ndk[(size_t)d * c->K + k] = t[(size_t)d * c->Kp + c->pos(k)];
    }
    if (nkw && c->V) {
        std::vector<int32_t> t((size_t)c->V * c->Kp);
        HIP_TRY(hipMemcpyAsync(t.data(), c->d_nkw.p, t.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int v = 0; v < c->V; ++v)
            for (int k = 0; k < c->K; ++k) nkw[(size_t)v * c->K + k] = t[(size_t)v * c->Kp + c->pos(k)];
    }
    if (nk) {
        std::vector<int32_t> t(c->Kp);
        HIP_TRY(hipMemcpyAsync(t.data(), c->d_nk.p, t.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        for (int k = 0; k < c->K; ++k) nk[k] = t[c->pos(k)];
    }
    HIP_TRY(hipStreamSynchronize(s));
    return LDA_OK;
}

int lda_get_dims(lda_ctx* c, int64_t dims[6]) {
    if (!c || !dims) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    dims[0] = c->K; dims[1] = c->V; dims[2] = c->D; dims[3] = (int64_t)c->N; dims[4] = c->Kp; dims[5] = c->kpl;
    return LDA_OK;
}

int lda_sample_frozen(lda_ctx* c, int64_t n, const int64_t* doc, const uint32_t* word, const uint16_t* z_old,
                      const float* u, uint16_t* z_new, float* total) {
    if (c && c->tw != LDA_TW_ONE) return fail(LDA_ERR_UNSUPPORTED, "lda_sample_frozen takes integer counts (TermWeight ONE)");
    if (c && c->has_priors) return fail(LDA_ERR_UNSUPPORTED, "lda_sample_frozen uses the symmetric eta");
    if (!c) return fail(LDA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (n < 0 || (n && (!doc || !word || !z_old || !u || !z_new))) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!c->has_corpus) return fail(LDA_ERR_STATE, "no corpus uploaded");
    for (int64_t t = 0; t < n; ++t) {
        if (doc[t] < 0 || doc[t] >= c->D) return fail(LDA_ERR_INVALID_ARGUMENT, "doc out of range at %lld", (long long)t);
        if (word[t] >= (uint32_t)c->V) return fail(LDA_ERR_INVALID_ARGUMENT, "word out of range at %lld", (long long)t);
        if (z_old[t] >= (uint16_t)c->K) return fail(LDA_ERR_INVALID_ARGUMENT, "z out of range at %lld", (long long)t);
        if (!(u[t] >= 0.0f && u[t] < 1.0f)) return fail(LDA_ERR_INVALID_ARGUMENT, "u out of [0,1) at %lld", (long long)t);
    }
    if (!n) return LDA_OK;
    HIP_TRY(hipSetDevice(c->device));
    const hipStream_t s = c->stream;
    DevBuf<int64_t> dd;
    DevBuf<uint32_t> dw;
    DevBuf<uint16_t> dz, dzn;
    DevBuf<float> du, dt;
    DevBuf<int32_t> dsc;
    int rc = LDA_OK;
    // K above 2048: three rows of Kp counts per workgroup
    const int bk_grid = (int)std::min<int64_t>(n, 1024);
    if (dd.alloc(n) || dw.alloc(n) || dz.alloc(n) || dzn.alloc(n) || du.alloc(n) || dt.alloc(n) ||
        ((c->kpl > lda::kSmallMaxKpl || c->exact_counts()) && dsc.alloc((size_t)bk_grid * 3 * c->Kp)))
        rc = fail(LDA_ERR_OUT_OF_MEMORY, "allocation failed");
    if (!rc && (hipMemcpyAsync(dd.p, doc, n * 8, hipMemcpyHostToDevice, s) || hipMemcpyAsync(dw.p, word, n * 4, hipMemcpyHostToDevice, s) ||
                hipMemcpyAsync(dz.p, z_old, n * 2, hipMemcpyHostToDevice, s) || hipMemcpyAsync(du.p, u, n * 4, hipMemcpyHostToDevice, s)))
        rc = fail(LDA_ERR_DEVICE, "upload failed");
    if (!rc) {
        hipError_t e = lda::launch_sample_frozen(c->kpl, c->d_ndk.p, c->d_nkw.p, c->d_nk.p, c->d_alpha.p, c->eta, c->veta, c->K,
                                                 c->V, c->Kp, n, dd.p, dw.p, dz.p, du.p, dzn.p, dt.p, dsc.p, bk_grid, s);
        if (e != hipSuccess) rc = fail(LDA_ERR_DEVICE, "sample_frozen: %s", hipGetErrorString(e));
    }
    if (!rc && (hipMemcpyAsync(z_new, dzn.p, n * 2, hipMemcpyDeviceToHost, s) ||
                (total && hipMemcpyAsync(total, dt.p, n * 4, hipMemcpyDeviceToHost, s)) || hipStreamSynchronize(s)))
        rc = fail(LDA_ERR_DEVICE, "download failed");
    dd.release(); dw.release(); dz.release(); dzn.release(); du.release(); dt.release(); dsc.release();
    return rc;
}

// Argument checks shared by the fold-in entry points; *done = true when there
// is nothing to do (D == 0).
static int check_infer_args(lda_ctx* c, const uint32_t* words, const uint64_t* doc_off, int64_t D, int iterations,
                            int mode, float* theta, double* ll, bool* done) {
    *done = false;
    if (!c) return fail(LDA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (D < 0 || (D && (!doc_off || !theta || !ll))) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    if (iterations < 0) return fail(LDA_ERR_INVALID_ARGUMENT, "iterations must be >= 0");
    if (mode != 0 && mode != 1) return fail(LDA_ERR_INVALID_ARGUMENT, "mode must be 0 or 1");
    if (!D) { *done = true; return LDA_OK; }
    if (D > 0x7FFFFFFFll) return fail(LDA_ERR_INVALID_ARGUMENT, "too many documents");
    if (doc_off[0] != 0) return fail(LDA_ERR_INVALID_ARGUMENT, "doc_off[0] must be 0");
    if (doc_off[D] && !words) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    for (int64_t d = 0; d < D; ++d) {
        if (doc_off[d + 1] < doc_off[d]) return fail(LDA_ERR_INVALID_ARGUMENT, "doc_off must be non-decreasing");
    }
    std::atomic<int64_t> unsorted(INT64_MAX);
    host_parallel(D, 1024, [&](int64_t lo, int64_t hi) {
        for (int64_t d = lo; d < hi; ++d)
            for (uint64_t g = doc_off[d] + 1; g < doc_off[d + 1]; ++g)
                if (words[g] < words[g - 1]) { atomic_min_i64(unsorted, d); return; }
    });
    if (unsorted.load() != INT64_MAX)
        return fail(LDA_ERR_INVALID_ARGUMENT, "words must be sorted within each document (document %lld)",
                    (long long)unsorted.load());
    return LDA_OK;
}

// Segment w's first token of an n-token document split into S segments (the
// fold-in kernels' rule): floor(n*w/S), moved forward to a word-run start.
static uint64_t seg_start_host(const uint32_t* w, uint64_t n, int i, int S) {
    if (i <= 0) return 0;
    if (i >= S) return n;
    uint64_t x = n * (uint64_t)i / (uint64_t)S;
    if (x == 0) return 0;
    while (x < n && w[x] == w[x - 1]) ++x;
    return x;
}

// Fold-in of term-weighted models and of models with word priors: each run
// of equal word ids with a private column; documents of more than seg_tokens
// tokens split into S segments (the integer fold-in's rule and options).
// Mode 0: jobs grouped by S, longest first, one persistent launch per
// segment count on the fold-in streams, each followed on its stream by the
// finalize kernel of its documents; mode 1: one launch per document.  Arguments checked by check_infer_args.
}  // extern "C"

static int infer_general_batch(lda_ctx* c, lda_host::GenFold& B, hipStream_t s, lda_host::FoldStreams& fs,
                               hipEvent_t e0, hipEvent_t e1, const uint32_t* words, const uint64_t* doc_off, int64_t D,
                               const float* weights, int iterations, uint64_t infer_seed, int mode, float* theta,
                               double* ll, uint16_t* z_out, float* ms_out);

// Private columns of one fold-in launch: batches whose runs need more are
// folded in as consecutive groups of documents (each document's result does
// not depend on the others in its batch).
constexpr double kGenCellBudget = 8e9;

int lda_host::infer_general_on(lda_ctx* c, lda_host::GenFold& B, hipStream_t s, lda_host::FoldStreams& fs,
                               hipEvent_t e0, hipEvent_t e1, const uint32_t* words, const uint64_t* doc_off, int64_t D,
                               const float* weights, int iterations, uint64_t infer_seed, int mode, float* theta,
                               double* ll, uint16_t* z_out, float* ms_out) {
    const uint32_t Vu = (uint32_t)c->V;
    std::vector<int64_t> runs(D, 0);
    host_parallel(D, 256, [&](int64_t lo, int64_t hi) {
        for (int64_t d = lo; d < hi; ++d) {
            int64_t nr = 0, prev = -1;
            for (uint64_t g = doc_off[d]; g < doc_off[d + 1]; ++g) {
                const uint32_t v = words[g];
                if (v >= Vu || (int64_t)v == prev) continue;
                ++nr;
                prev = v;
            }
            runs[d] = nr;
        }
    });
    const double per_run = (double)c->Kp * 4.0;
    float total_ms = 0.0f;
    std::vector<uint64_t> off;
    for (int64_t g0 = 0; g0 < D;) {
        int64_t g1 = g0 + 1;
        double bytes = (double)runs[g0] * per_run;
        while (g1 < D && bytes + (double)runs[g1] * per_run <= kGenCellBudget) bytes += (double)runs[g1++] * per_run;
        if (g0 == 0 && g1 == D) return infer_general_batch(c, B, s, fs, e0, e1, words, doc_off, D, weights, iterations,
                                                          infer_seed, mode, theta, ll, z_out, ms_out);
        const uint64_t base = doc_off[g0];
        off.resize((size_t)(g1 - g0) + 1);
        for (int64_t d = g0; d <= g1; ++d) off[(size_t)(d - g0)] = doc_off[d] - base;
        float ms = 0.0f;
        const int rc = infer_general_batch(c, B, s, fs, e0, e1, words + base, off.data(), g1 - g0,
                                           weights ? weights + base : nullptr, iterations, infer_seed, mode,
                                           theta + (size_t)g0 * c->K, ll + g0, z_out ? z_out + base : nullptr, &ms);
        if (rc) return rc;
        total_ms += ms;
        g0 = g1;
    }
    if (ms_out) *ms_out = total_ms;
    return LDA_OK;
}

static int infer_general_batch(lda_ctx* c, lda_host::GenFold& B, hipStream_t s, lda_host::FoldStreams& fs,
                               hipEvent_t e0, hipEvent_t e1, const uint32_t* words, const uint64_t* doc_off, int64_t D,
                               const float* weights, int iterations, uint64_t infer_seed, int mode, float* theta,
                               double* ll, uint16_t* z_out, float* ms_out) {
    const bool wt = c->tw != LDA_TW_ONE;
    const uint64_t N = doc_off[D];
    constexpr int RW = lda::kMaxFoldSegments + 1;
    // integer counts with a document over 2^20 tokens or an alpha >= 2^20:
    // the exact-count kernels (one wavefront per document)
    bool exact = false;
    if (!wt) {
        exact = c->exact_alpha;
        for (int64_t d = 0; d < D && !exact; ++d) exact = doc_off[d + 1] - doc_off[d] > lda_host::kMaxDocTokens;
    }
    int per_cu[lda::kFoldBuckets] = {};
    int gen_max = 1;
    for (int bk = 0; bk < lda::kFoldBuckets; ++bk) {
        const int sg = 1 << bk;
        if (exact) per_cu[bk] = sg == 1 ? std::max(lda::bk_infer_gen_blocks_per_cu(false, c->has_priors, c->kpl), 1) : 0;
        else per_cu[bk] = sg <= c->seg_max ? lda::infer_gen_blocks_per_cu(c->kpl, wt, c->has_priors, sg) : 0;
        if (per_cu[bk] <= 0) break;
        gen_max = sg;
    }
    if (per_cu[0] <= 0) return fail(LDA_ERR_DEVICE, "fold-in kernel does not fit the device for K = %d", c->K);
    std::vector<int> segs(D);
    for (int64_t d = 0; d < D; ++d) segs[d] = lda::fold_segments(doc_off[d + 1] - doc_off[d], c->seg_tokens, gen_max);
    std::vector<uint32_t> order(D);
    std::iota(order.begin(), order.end(), 0u);
    if (mode == 0)
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
            if (segs[a] != segs[b]) return segs[a] < segs[b];
            return (doc_off[a + 1] - doc_off[a]) > (doc_off[b + 1] - doc_off[b]);
        });
    std::vector<lda::InferJob> jobs(D);
    std::vector<int64_t> job_run(D + 1), seg_r((size_t)D * RW, 0);
    std::vector<uint64_t> seg_b((size_t)D * RW, 0);
    // runs of equal valid word ids: counted per job, then written from each
    // job's first run index (both passes over the jobs on the host threads)
    const uint32_t Vu = (uint32_t)c->V;
    host_parallel(D, 256, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            const uint32_t d = order[i];
            jobs[i] = lda::InferJob{doc_off[d], doc_off[d + 1], 0, iterations, iterations, (int32_t)d};
            int64_t nr = 0, prev = -1;
            for (uint64_t g = doc_off[d]; g < doc_off[d + 1]; ++g) {
                const uint32_t v = words[g];
                if (v >= Vu || (int64_t)v == prev) continue;
                ++nr;
                prev = v;
            }
            job_run[i + 1] = nr;
        }
    });
    job_run[0] = 0;
    for (int64_t i = 0; i < D; ++i) job_run[i + 1] += job_run[i];
    const size_t R = (size_t)job_run[D];
    std::vector<uint32_t> run_word(R);
    host_parallel(D, 256, [&](int64_t lo, int64_t hi) {
        std::vector<uint64_t> run_tok;  // first token of each run of the document (relative)
        for (int64_t i = lo; i < hi; ++i) {
            const uint32_t d = order[i];
            const uint64_t b = doc_off[d], n = doc_off[d + 1] - b;
            run_tok.clear();
            int64_t r = job_run[i], prev = -1;
            for (uint64_t g = b; g < b + n; ++g) {
                const uint32_t v = words[g];
                if (v >= Vu || (int64_t)v == prev) continue;
                run_word[(size_t)r++] = v;
                run_tok.push_back(g - b);
                prev = v;
            }
            const int S = segs[d];
            for (int w = 0; w <= S; ++w) {
                const uint64_t x = seg_start_host(words + b, n, w, S);
                seg_b[(size_t)i * RW + w] = b + x;
                seg_r[(size_t)i * RW + w] =
                    job_run[i] + (int64_t)(std::lower_bound(run_tok.begin(), run_tok.end(), x) - run_tok.begin());
            }
        }
    });
    // the finalize's chunks: at most kGenChunkRuns runs of one document each
    std::vector<int64_t> chunk_base(D + 1), chunk_rb;
    for (int64_t i = 0; i < D; ++i) {
        chunk_base[i] = (int64_t)chunk_rb.size();
        for (int64_t r = job_run[i]; r < job_run[i + 1]; r += lda::kGenChunkRuns) chunk_rb.push_back(r);
    }
    chunk_base[D] = (int64_t)chunk_rb.size();
    const size_t NC = chunk_rb.size();
    chunk_rb.push_back(job_run[D]);
    const double cell_bytes = (double)R * c->Kp * 4.0;
    if (cell_bytes > 64e9)
        return fail(LDA_ERR_OUT_OF_MEMORY, "a document needs %.1f GB of private columns", cell_bytes / 1e9);
    int64_t lo[lda::kFoldBuckets] = {}, hi[lda::kFoldBuckets] = {};
    if (mode == 0) {
        int64_t i = 0;
        for (int bk = 0; bk < lda::kFoldBuckets; ++bk) {
            lo[bk] = i;
            while (i < D && lda::fold_bucket(segs[order[i]]) == bk) ++i;
            hi[bk] = i;
        }
    }
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(B.words.alloc(N));
    HIP_TRY(B.z.alloc(N));
    HIP_TRY(B.jobs.alloc(D));
    HIP_TRY(B.counter.alloc(lda::kFoldBuckets));
    HIP_TRY(B.theta.alloc((size_t)D * c->K));
    HIP_TRY(B.ll.alloc(D));
    if (wt) HIP_TRY(B.wt.alloc(N));
    HIP_TRY(B.jobrun.alloc(D + 1));
    HIP_TRY(B.segr.alloc((size_t)D * RW));
    HIP_TRY(B.segb.alloc((size_t)D * RW));
    HIP_TRY(B.runword.alloc(R));
    HIP_TRY(B.cells.alloc(R * (size_t)c->Kp));
    HIP_TRY(B.ndk.alloc((size_t)D * c->Kp));
    HIP_TRY(B.nk.alloc((size_t)D * c->Kp));
    HIP_TRY(B.docw.alloc(D));
    HIP_TRY(B.chunkbase.alloc(D + 1));
    HIP_TRY(B.chunkrb.alloc(NC + 1));
    HIP_TRY(B.chunkll.alloc(NC ? NC : 1));
    if (N) HIP_TRY(hipMemcpyAsync(B.words.p, words, N * 4, hipMemcpyHostToDevice, s));
    if (wt && N) HIP_TRY(hipMemcpyAsync(B.wt.p, weights, N * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(B.jobs.p, jobs.data(), D * sizeof(lda::InferJob), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(B.jobrun.p, job_run.data(), (D + 1) * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(B.segr.p, seg_r.data(), seg_r.size() * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(B.chunkbase.p, chunk_base.data(), (D + 1) * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(B.chunkrb.p, chunk_rb.data(), (NC + 1) * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(B.segb.p, seg_b.data(), seg_b.size() * 8, hipMemcpyHostToDevice, s));
    if (R) HIP_TRY(hipMemcpyAsync(B.runword.p, run_word.data(), R * 4, hipMemcpyHostToDevice, s));
    lda::InferGenParams p{};
    p.exact = exact ? 1 : 0;
    p.words = B.words.p; p.z = B.z.p; p.wt = wt ? B.wt.p : nullptr;
    p.jobs = B.jobs.p; p.njobs = D;
    p.job_run = B.jobrun.p; p.seg_b = B.segb.p; p.seg_r = B.segr.p;
    p.chunk_base = B.chunkbase.p; p.chunk_rb = B.chunkrb.p; p.chunk_ll = B.chunkll.p;
    p.run_word = B.runword.p; p.cells = B.cells.p;
    p.ndk_out = B.ndk.p; p.nk_out = B.nk.p; p.docw = B.docw.p;
    p.nkw = c->d_nkw.p; p.nk = c->d_nk.p; p.alpha = c->d_alpha.p; p.alpha_cdf = c->d_alpha_cdf.p;
    p.eta = c->eta; p.veta = c->veta;
    p.eta_d = (double)c->eta; p.veta_d = (double)c->V * (double)c->eta;
    p.K = c->K; p.V = c->V; p.Kp = c->Kp;
    if (c->has_priors) {
        p.prior_row = c->d_prior_row.p; p.prior_eta = c->d_prior_eta.p;
        p.etasum = c->d_etasum.p; p.etasum_d = c->d_etasum_d.p;
    }
    p.seed_lo = (uint32_t)infer_seed; p.seed_hi = (uint32_t)(infer_seed >> 32);
    p.theta = B.theta.p; p.ll = B.ll.p;
    double asd = 0.0;
    float asf = 0.0f;
    for (int k = 0; k < c->K; ++k) { asd += (double)c->alpha[k]; asf = asf + c->alpha[k]; }
    p.asum_d = asd; p.asum_f = asf;
    HIP_TRY(hipEventRecord(e0, s));
    if (mode == 0) {
        HIP_TRY(lda::launch_fill_u32(B.counter.p, 0u, lda::kFoldBuckets, s));
        HIP_TRY(hipEventRecord(fs.fork, s));
        // the longest documents (most segments) first, each segment count on its own stream
        for (int bk = lda::kFoldBuckets - 1; bk >= 0; --bk) {
            if (hi[bk] <= lo[bk]) continue;
            lda::InferGenParams pb = p;
            pb.jobs += lo[bk];
            pb.job_run += lo[bk];
            pb.seg_b += lo[bk] * RW;
            pb.seg_r += lo[bk] * RW;
            pb.chunk_base += lo[bk];
            pb.njobs = hi[bk] - lo[bk];
            pb.counter = B.counter.p + bk;
            const int grid = (int)std::min<int64_t>((int64_t)c->num_cu * per_cu[bk], pb.njobs);
            const hipStream_t sb = bk == 0 ? s : fs.aux[bk - 1];
            if (bk) HIP_TRY(hipStreamWaitEvent(sb, fs.fork, 0));
            HIP_TRY(lda::launch_infer_gen(c->kpl, 1 << bk, pb, grid, false, sb));
            // the bucket's theta and ll as soon as its documents are done, beside the other buckets' sampling
            const int64_t nch = chunk_base[hi[bk]] - chunk_base[lo[bk]];
            const int fgrid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)c->num_cu * 16, std::max(nch, pb.njobs)));
            HIP_TRY(lda::launch_infer_gen(c->kpl, 1, pb, fgrid, true, sb));
            if (bk) HIP_TRY(hipEventRecord(fs.join[bk - 1], sb));
        }
        for (int bk = 1; bk < lda::kFoldBuckets; ++bk)
            if (hi[bk] > lo[bk]) HIP_TRY(hipStreamWaitEvent(s, fs.join[bk - 1], 0));
    } else {
        p.counter = nullptr;
        for (int64_t d = 0; d < D; ++d) {
            p.job_base = d;
            HIP_TRY(lda::launch_infer_gen(c->kpl, segs[d], p, 1, false, s));
            HIP_TRY(lda::launch_infer_gen(c->kpl, 1, p, 1, true, s));
            HIP_TRY(hipStreamSynchronize(s));
        }
    }
    HIP_TRY(hipEventRecord(e1, s));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    if (ms_out) *ms_out = ms;
    HIP_TRY(hipMemcpyAsync(theta, B.theta.p, (size_t)D * c->K * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(ll, B.ll.p, D * 8, hipMemcpyDeviceToHost, s));
    if (z_out && N) HIP_TRY(hipMemcpyAsync(z_out, B.z.p, N * 2, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return LDA_OK;
}

extern "C" {

static int infer_general(lda_ctx* c, const uint32_t* words, const uint64_t* doc_off, int64_t D, const float* weights,
                         int iterations, uint64_t infer_seed, int mode, float* theta, double* ll, uint16_t* z_out) {
    c->stat_ms = c->stat_launches = c->stat_tokens = 0.0;
    float ms = 0.0f;
    const int rc = lda_host::infer_general_on(c, c->gen, c->stream, c->fold_streams, c->ev0, c->ev1, words, doc_off, D,
                                              weights, iterations, infer_seed, mode, theta, ll, z_out, &ms);
    if (rc) return rc;
    c->stat_ms = ms;
    c->stat_launches = mode == 0 ? 1.0 : (double)D;
    c->stat_tokens = (double)doc_off[D] * iterations;
    return LDA_OK;
}

int lda_infer_w(lda_ctx* c, const uint32_t* words, const uint64_t* doc_off, int64_t D, const float* weights,
                int iterations, uint64_t infer_seed, int mode, float* theta, double* ll, uint16_t* z_out) {
    bool done = false;
    int rc = check_infer_args(c, words, doc_off, D, iterations, mode, theta, ll, &done);
    if (rc || done) return rc;
    if (c->tw == LDA_TW_ONE) {
        if (weights) return fail(LDA_ERR_INVALID_ARGUMENT, "TermWeight ONE models take no token weights");
        return lda_infer(c, words, doc_off, D, iterations, infer_seed, mode, theta, ll, z_out);
    }
    if (!weights && doc_off[D]) return fail(LDA_ERR_INVALID_ARGUMENT, "term-weighted models need the token weights");
    std::atomic<int64_t> bad(INT64_MAX);
    host_parallel((int64_t)doc_off[D], 1 << 16, [&](int64_t lo, int64_t hi) {
        for (int64_t g = lo; g < hi; ++g)
            if (words[g] < (uint32_t)c->V && !(weights[g] >= 0.0f && weights[g] < 1e30f)) {
                atomic_min_i64(bad, g);
                return;
            }
    });
    if (bad.load() != INT64_MAX)
        return fail(LDA_ERR_INVALID_ARGUMENT, "token weight %llu is negative or not finite", (unsigned long long)bad.load());
    return infer_general(c, words, doc_off, D, weights, iterations, infer_seed, mode, theta, ll, z_out);
}

int lda_infer(lda_ctx* c, const uint32_t* words, const uint64_t* doc_off, int64_t D, int iterations, uint64_t infer_seed,
              int mode, float* theta, double* ll, uint16_t* z_out) {
    bool done = false;
    int rc = check_infer_args(c, words, doc_off, D, iterations, mode, theta, ll, &done);
    if (rc || done) return rc;
    if (c->tw != LDA_TW_ONE)
        return fail(LDA_ERR_INVALID_ARGUMENT, "term-weighted models need the token weights: call lda_infer_w");
    // K above 2048: the run-column fold-in (one wavefront per document), the
    // same draws as the register-resident sampler for integer counts
    bool run_column = c->kpl > lda::kSmallMaxKpl || c->exact_alpha;
    for (int64_t d = 0; d < D && !run_column; ++d) run_column = doc_off[d + 1] - doc_off[d] > kMaxDocTokens;
    if (run_column)
        return infer_general(c, words, doc_off, D, nullptr, iterations, infer_seed, mode, theta, ll, z_out);
    const uint64_t N = doc_off[D];
    HIP_TRY(hipSetDevice(c->device));
    const hipStream_t s = c->stream;
    HIP_TRY(c->i_words.alloc(N));
    HIP_TRY(c->i_z.alloc(N));
    HIP_TRY(c->i_jobs.alloc(D));
    HIP_TRY(c->i_params.alloc(1));
    HIP_TRY(c->i_counter.alloc(lda::kFoldBuckets));
    HIP_TRY(c->i_theta.alloc((size_t)D * c->K));
    HIP_TRY(c->i_ll.alloc(D));
    const FoldArena a{c->i_words.p, c->i_z.p, c->i_jobs.p, c->i_params.p, c->i_counter.p, c->i_theta.p, c->i_ll.p};
    if (N) HIP_TRY(hipMemcpyAsync(a.words, words, N * 4, hipMemcpyHostToDevice, s));
    std::vector<int> segs;
    std::vector<lda::InferJob> jobs;
    lda::InferParams p;
    build_fold_jobs(c, doc_off, D, iterations, infer_seed, mode, a, segs, jobs, p);
    HIP_TRY(hipMemcpyAsync(a.jobs, jobs.data(), D * sizeof(lda::InferJob), hipMemcpyHostToDevice, s));
    c->stat_ms = c->stat_launches = c->stat_tokens = 0.0;
    if (mode == 0) {
        HIP_TRY(hipMemcpyAsync(a.params, &p, sizeof p, hipMemcpyHostToDevice, s));
        // the graph launches the buckets up to seg_max, each with its full persistent grid
        if (!c->infer_exec || c->infer_exec_segmax != c->seg_max) {
            if (c->infer_exec) { (void)hipGraphExecDestroy(c->infer_exec); c->infer_exec = nullptr; }
            int grids[lda::kFoldBuckets];
            for (int bk = 0; bk < lda::kFoldBuckets; ++bk) grids[bk] = (1 << bk) <= c->seg_max ? c->fold_grid[bk] : 0;
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            hipError_t e1 = lda::launch_fill_u32(a.counters, 0u, lda::kFoldBuckets, s);
            hipError_t e2 = e1 == hipSuccess ? lda_host::launch_fold_buckets(c->kpl, c->has_priors, a.params, grids, s,
                                                                             c->fold_streams)
                                             : e1;
            if (e2 == hipSuccess) e2 = lda::launch_infer_finalize(c->kpl, c->has_priors, a.params, c->fold_grid[0], s);
            hipError_t e3 = hipStreamEndCapture(s, &g);
            if (e2 != hipSuccess || e3 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return fail(LDA_ERR_DEVICE, "fold-in graph capture failed: %s", hipGetErrorString(e2 != hipSuccess ? e2 : e3));
            }
            hipError_t e4 = hipGraphInstantiate(&c->infer_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e4 != hipSuccess) return fail(LDA_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e4));
            c->infer_exec_segmax = c->seg_max;
        }
        HIP_TRY(hipEventRecord(c->ev0, s));
        HIP_TRY(hipGraphLaunch(c->infer_exec, s));
        HIP_TRY(hipEventRecord(c->ev1, s));
        c->stat_launches = 1;
    } else {
        // plain: one launch and one synchronisation per document
        HIP_TRY(hipEventRecord(c->ev0, s));
        for (int64_t d = 0; d < D; ++d) {
            p.job_base = d;
            HIP_TRY(hipMemcpyAsync(a.params, &p, sizeof p, hipMemcpyHostToDevice, s));
            HIP_TRY(lda::launch_infer(c->kpl, segs[d], c->has_priors, a.params, 1, s));
            HIP_TRY(lda::launch_infer_finalize(c->kpl, c->has_priors, a.params, 1, s));
            HIP_TRY(hipStreamSynchronize(s));
        }
        HIP_TRY(hipEventRecord(c->ev1, s));
        c->stat_launches = (double)D;
    }
    return fold_download(c, a, D, N, iterations, theta, ll, z_out);
}

int lda_infer_noalloc(lda_ctx* c, const uint32_t* words, const uint64_t* doc_off, int64_t D, int iterations,
                      uint64_t infer_seed, float* theta, double* ll, uint16_t* z_out) {
    bool done = false;
    int rc = check_infer_args(c, words, doc_off, D, iterations, 0, theta, ll, &done);
    if (rc || done) return rc;
    if (c->tw != LDA_TW_ONE)
        return fail(LDA_ERR_INVALID_ARGUMENT, "term-weighted models need the token weights: call lda_infer_w");
    if (c->kpl > lda::kSmallMaxKpl)
        return fail(LDA_ERR_UNSUPPORTED, "lda_infer_noalloc takes K <= 2048 (K = %d: use lda_infer)", c->K);
    if (c->exact_alpha)
        return fail(LDA_ERR_UNSUPPORTED, "lda_infer_noalloc takes alpha below 2^20 (use lda_infer)");
    for (int64_t d = 0; d < D; ++d)
        if (doc_off[d + 1] - doc_off[d] > kMaxDocTokens)
            return fail(LDA_ERR_UNSUPPORTED, "lda_infer_noalloc takes documents of at most 2^20 tokens (use lda_infer)");
    const uint64_t N = doc_off[D];
    // the arena inside the scratch array, each region 256-byte aligned
    const size_t o_params = 0;
    const size_t o_counters = o_params + align256(sizeof(lda::InferParams));
    const size_t o_jobs = o_counters + align256(lda::kFoldBuckets * sizeof(uint32_t));
    const size_t o_words = o_jobs + align256((size_t)D * sizeof(lda::InferJob));
    const size_t o_z = o_words + align256((size_t)N * sizeof(uint32_t));
    const size_t o_theta = o_z + align256((size_t)N * sizeof(uint16_t));
    const size_t o_ll = o_theta + align256((size_t)D * c->K * sizeof(float));
    const size_t need = o_ll + align256((size_t)D * sizeof(double));
    if (need > kNoAllocScratchBytes)
        return fail(LDA_ERR_INVALID_ARGUMENT,
                    "lda_infer_noalloc: %lld documents of %llu tokens at K = %d need %.1f MiB of scratch; the limit is "
                    "%zu MiB (use lda_infer)",
                    (long long)D, (unsigned long long)N, c->K, need / 1048576.0, kNoAllocScratchBytes >> 20);
    unsigned char* base;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = static_cast<unsigned char*>(g_scratch_dev[c->device]);
    }
    if (!base) return fail(LDA_ERR_STATE, "the fold-in scratch of device %d is not set up", c->device);
    HIP_TRY(hipSetDevice(c->device));
    const hipStream_t s = c->stream;
    const FoldArena a{reinterpret_cast<uint32_t*>(base + o_words), reinterpret_cast<uint16_t*>(base + o_z),
                      reinterpret_cast<lda::InferJob*>(base + o_jobs), reinterpret_cast<lda::InferParams*>(base + o_params),
                      reinterpret_cast<uint32_t*>(base + o_counters), reinterpret_cast<float*>(base + o_theta),
                      reinterpret_cast<double*>(base + o_ll)};
    std::vector<int> segs;
    std::vector<lda::InferJob> jobs;
    lda::InferParams p;
    build_fold_jobs(c, doc_off, D, iterations, infer_seed, 0, a, segs, jobs, p);
    if (N) HIP_TRY(hipMemcpyAsync(a.words, words, N * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(a.jobs, jobs.data(), D * sizeof(lda::InferJob), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(a.params, &p, sizeof p, hipMemcpyHostToDevice, s));
    c->stat_ms = c->stat_launches = c->stat_tokens = 0.0;
    int grids[lda::kFoldBuckets];
    for (int bk = 0; bk < lda::kFoldBuckets; ++bk) grids[bk] = (1 << bk) <= c->seg_max ? c->fold_grid[bk] : 0;
    HIP_TRY(hipEventRecord(c->ev0, s));
    // the bucket counters start at zero (a fill kernel: the previous call's
    // values are still there)
    HIP_TRY(lda::launch_fill_u32(a.counters, 0u, lda::kFoldBuckets, s));
    HIP_TRY(lda_host::launch_fold_buckets(c->kpl, c->has_priors, a.params, grids, s, c->fold_streams));
    HIP_TRY(lda::launch_infer_finalize(c->kpl, c->has_priors, a.params, c->fold_grid[0], s));
    HIP_TRY(hipEventRecord(c->ev1, s));
    c->stat_launches = 1;
    return fold_download(c, a, D, N, iterations, theta, ll, z_out);
}

int lda_set_infer_options(lda_ctx* c, uint32_t seg_tokens, int seg_max) {
    if (!c) return fail(LDA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (seg_max != 1 && seg_max != 2 && seg_max != 4 && seg_max != 8)
        return fail(LDA_ERR_INVALID_ARGUMENT, "seg_max must be 1, 2, 4 or 8");
    if (seg_max > c->seg_supported)
        return fail(LDA_ERR_UNSUPPORTED, "seg_max %d exceeds the %d wavefronts per document the fold-in kernel fits at K = %d",
                    seg_max, c->seg_supported, c->K);
    c->seg_tokens = seg_tokens;
    c->seg_max = seg_max;
    return LDA_OK;
}

int lda_get_infer_options(lda_ctx* c, uint32_t* seg_tokens, int* seg_max, int* seg_supported) {
    if (!c) return fail(LDA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (seg_tokens) *seg_tokens = c->seg_tokens;
    if (seg_max) *seg_max = c->seg_max;
    if (seg_supported) *seg_supported = c->seg_supported;
    return LDA_OK;
}

int lda_kernel_stats(lda_ctx* c, double stats[4]) {
    if (!c || !stats) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    stats[0] = c->stat_ms;
    stats[1] = c->stat_launches;
    stats[2] = c->stat_tokens;
    stats[3] = c->narrow_used ? 1.0 : 0.0;
    return LDA_OK;
}

int lda_set_narrow_columns(lda_ctx* c, int enable) {
    if (!c) return fail(LDA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    c->narrow_enabled = enable != 0;
    return LDA_OK;
}

int lda_set_atomic_group(lda_ctx* c, int tokens) {
    if (!c) return fail(LDA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (tokens < 0 || tokens > 64 || (tokens & (tokens - 1)) != 0)
        return fail(LDA_ERR_INVALID_ARGUMENT, "tokens must be 0 or a power of two up to 64 (got %d)", tokens);
    c->dev_col_group = tokens;
    return LDA_OK;
}

}  // extern "C"
