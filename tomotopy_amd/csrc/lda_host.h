// Host-core internals shared by the C-ABI translation units (lda_capi.cpp,
// lda_serve.cpp): error reporting, device buffers and the model context.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/tomotopy_amd.h"
#include "lda_internal.h"

namespace lda_host {

// Sets the thread-local message returned by lda_last_error() and returns code.
int fail(int code, const char* fmt, ...);

#define HIP_TRY(expr)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return ::lda_host::fail(e_ == hipErrorOutOfMemory ? LDA_ERR_OUT_OF_MEMORY : LDA_ERR_DEVICE, \
                                    "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count) {
        if (count <= n && p) return hipSuccess;
        release();
        size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        hipError_t e = hipMalloc((void**)&p, bytes);
        if (e == hipSuccess) n = count;
        else p = nullptr;
        return e;
    }
};

// The K <= 2048 samplers keep float(n_dk) + alpha_k and recover n_dk by
// rounding, which is exact while n_dk + alpha_k < 2^21.  Integer counts with a
// document over kMaxDocTokens tokens or an alpha >= kMaxAlpha sample on the
// exact-count kernels instead (SweepParams::exact).
constexpr uint64_t kMaxDocTokens = 1u << 20;
constexpr float kMaxAlpha = (float)(1u << 20);

// Streams and events to run the fold-in buckets of one launch concurrently:
// bucket 0 on the caller's stream, bucket b > 0 on aux[b-1], forked from and
// joined back into the caller's stream (also under graph capture).
struct FoldStreams {
    hipStream_t aux[lda::kFoldBuckets - 1] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[lda::kFoldBuckets - 1] = {};
    hipError_t create();
    void destroy();
};

// Launches the fold-in kernel for every bucket b with grids[b] > 0 (segments
// 1 << b per document) on the streams of fs, joined back into s.
hipError_t launch_fold_buckets(int kpl, bool asym, const lda::InferParams* p_dev, const int grids[lda::kFoldBuckets],
                               hipStream_t s, FoldStreams& fs);

// Device buffers of one run-column fold-in call (InferGenParams): the token
// arena, token weights, jobs, runs, private columns, per-document state and
// outputs.  A context holds one set; each server lane of a term-weighted
// model holds its own.
struct GenFold {
    DevBuf<uint32_t> words, counter, runword;
    DevBuf<uint16_t> z;
    DevBuf<lda::InferJob> jobs;
    DevBuf<float> theta, wt, docw;
    DevBuf<double> ll, chunkll;
    DevBuf<int64_t> jobrun, segr, chunkbase, chunkrb;
    DevBuf<uint64_t> segb;
    DevBuf<int32_t> cells, ndk, nk;
    void release() {
        words.release(); counter.release(); runword.release(); z.release(); jobs.release(); theta.release();
        wt.release(); docw.release(); ll.release(); chunkll.release(); jobrun.release(); segr.release();
        chunkbase.release(); chunkrb.release(); segb.release(); cells.release(); ndk.release(); nk.release();
    }
};

// The run-column fold-in of D documents (host arrays; arguments checked by
// the caller) on buffers B and stream s with bucket streams fs: mode 0 all
// documents batched, mode 1 one launch per document.  e0 / e1 time the
// kernels (*ms_out).  Used by lda_infer_w, lda_infer (K above 2048) and the
// server of term-weighted models.
int infer_general_on(lda_ctx* c, GenFold& B, hipStream_t s, FoldStreams& fs, hipEvent_t e0, hipEvent_t e1,
                     const uint32_t* words, const uint64_t* doc_off, int64_t D, const float* weights, int iterations,
                     uint64_t infer_seed, int mode, float* theta, double* ll, uint16_t* z_out, float* ms_out);

}  // namespace lda_host

using lda_host::DevBuf;

struct lda_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // device-wide sweeps: event pairs of the last two sweeps; a sweep's time
    // is read once the next one is queued, so the host does not leave the
    // device idle between sweeps (dw_pending: the pair not yet read, or -1)
    hipEvent_t dw_ev[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    int dw_next = 0, dw_pending = -1;
    int K = 0, V = 0, kpl = 1, Kp = 64;
    float eta = 0.01f, veta = 0.0f;
    int tw = LDA_TW_ONE;
    uint64_t seed = 0;
    std::vector<float> alpha;  // natural order
    uint64_t global_step = 0;
    uint32_t optim_interval = 10, burn_in = 0;
    int num_cu = 1;

    // an alpha reached 2^20 (set or optimised): integer counts then sample on
    // the exact-count kernels (SweepParams::exact)
    bool exact_alpha = false;
    bool exact_counts() const { return tw == LDA_TW_ONE && (exact_alpha || (uint64_t)max_len > lda_host::kMaxDocTokens); }

    // corpus
    int64_t D = 0;
    uint64_t N = 0;
    uint64_t n_valid = 0;
    std::vector<uint64_t> off;
    std::vector<uint32_t> words;
    bool has_corpus = false;
    // term-weighted models (tw != LDA_TW_ONE): float counts in the count
    // buffers, token weights, per-document weight sums and their total
    std::vector<float> wt;
    double weighted_n = 0.0;
    DevBuf<float> d_wt, d_docw, d_rowpart;
    // word priors (set_word_prior, LDAModel.hpp:724-737): prior_row[v] = row
    // of word v's eta column in prior_eta (rows of Kp, lane layout) or -1;
    // per-topic sums of eta_kv in float (sampler) and double (FP64 LL)
    bool has_priors = false;
    // partition chunks from these word frequencies instead of the corpus's
    // (fold-in together: the model's vocabCf, LDAModel.hpp:441-449)
    std::vector<uint64_t> chunk_cf;
    // term-weighted upload: the float counts start from these (lane layout,
    // V x Kp and Kp) instead of 0 when set (fold-in together: the model's)
    std::vector<float> base_nkw, base_nk;
    DevBuf<int32_t> d_prior_row;
    DevBuf<float> d_prior_eta, d_etasum;
    DevBuf<double> d_etasum_d;

    DevBuf<float> d_alpha, d_alpha_cdf;
    DevBuf<int32_t> d_nkw, d_nk, d_nk_next, d_ndk, d_doclen, d_nk_priv, d_nkw_loc;
    // K above 2048: the device-wide sweep's per-worker rows (private n_k and column copy)
    DevBuf<int32_t> d_bk_scratch;
    DevBuf<uint32_t> d_words, d_order, d_counter;
    DevBuf<uint16_t> d_z;
    DevBuf<uint64_t> d_off, d_coff;
    DevBuf<double> d_part;
    int part_W = 0;
    // alpha fixed point on the device: tomoto's digammaT table, {sum alpha,
    // denominator}, the out-of-range flag, the count histograms; the largest
    // document length in valid tokens
    DevBuf<float> d_dglut, d_aux;
    DevBuf<uint32_t> d_bad;
    DevBuf<int32_t> d_hist, d_hist_len;
    int64_t max_len = 0;

    // fold-in: segmentation (fold_segments), the largest segment count the
    // kernels fit, persistent grid per bucket and the bucket streams
    uint32_t seg_tokens = 0;
    int seg_max = 1;
    int seg_supported = 1;
    int fold_grid[lda::kFoldBuckets] = {};
    lda_host::FoldStreams fold_streams;

    // fold-in buffers
    DevBuf<uint32_t> i_words, i_counter;
    DevBuf<lda::InferJob> i_jobs;
    DevBuf<lda::InferParams> i_params;
    DevBuf<uint16_t> i_z;
    DevBuf<float> i_theta;
    DevBuf<double> i_ll;
    // the run-column fold-in (term-weighted models, K above 2048)
    lda_host::GenFold gen;
    // batched fold-in graph: counter reset + persistent kernel reading i_params
    hipGraphExec_t infer_exec = nullptr;
    int infer_exec_segmax = 0;  // the buckets the cached graph launches

    // device-wide schedule: dev_auto picks the concurrency from the corpus
    // (at most ceil(D / kInFlightDiv) documents in flight); otherwise at most
    // dev_workers workgroups (0: as many as are resident), each of dev_waves
    // wavefronts sharing one private n_k
    bool dev_auto = true;
    int dev_workers = 0;
    int dev_waves = lda::kMaxSweepWaves;
    // tokens per group of shared n_kw atomics in the device-wide sweeps
    // (0: chosen per launch from the corpus and the wavefront count)
    int dev_col_group = 0;
    // partition rounds of one sweep captured as a graph (rebuilt when the
    // parameters it was captured with change); the sweep number is read
    // from d_sweep so that one graph serves every sweep
    hipGraphExec_t part_exec = nullptr;
    lda::SweepParams part_key{};
    bool part_key_w = false;
    DevBuf<uint32_t> d_sweep;

    // narrow columns of the device-wide sweep (TermWeight ONE): words
    // [nkw_hot, V) have fewer than kNarrowCf tokens; during an lda_train /
    // lda_sweep call their counts live as u16 in d_nkw16 (narrow_active) and
    // are written back to d_nkw at the end of the call
    bool narrow_enabled = true;
    bool narrow_active = false;
    bool narrow_used = false;  // the last call ran on narrow columns
    int nkw_hot = 0;
    DevBuf<uint16_t> d_nkw16;

    // read-out staging (lda_topic_word_dist, lda_count_by_topic)
    DevBuf<float> r_out;
    DevBuf<uint64_t> r_cnt;

    // lda_debug_fail_at_step: the training iteration at this global step
    // reports LDA_ERR_TRAINING (-1: never); private copies never inherit it
    int64_t fail_at_step = -1;

    // kernel timing of the last call
    double stat_ms = 0.0;
    double stat_launches = 0.0;
    double stat_tokens = 0.0;

    int pos(int k) const { return lda::kvec_pos(kpl, k); }
};
