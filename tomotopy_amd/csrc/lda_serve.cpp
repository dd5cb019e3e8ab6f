// lda_serve.cpp -- continuous-batching fold-in server over a frozen model.
//
// tomotopy answers infer() calls one document set at a time (py_LDA.cpp:345-411
// -> LDAModel::_infer, LDAModel.hpp:920-953).  A server keeps the trained
// n_kw / n_k resident on the device and batches requests as they arrive:
//   * admission: requests are taken in FIFO order once their tokens fit the
//     token arena (first fit over a free list) and their documents fit the
//     active-document limit; both are sized from the HBM budget;
//   * continuous batching: every step is one persistent fold-in launch over
//     all admitted documents.  A document advances by up to step_iters
//     iterations, fewer when it is long (at most step_tokens
//     token-iterations per document per step), so one long document does not
//     hold the others back; requests that arrive during a step join at the
//     next one, and finished documents leave after the step that ends them;
//   * every draw is keyed by (infer seed, token position in the document,
//     iteration), and each launch rebuilds n_dk from the stored z, so a
//     document's result does not depend on the requests it shared the device
//     with nor on how its iterations were split into steps: it equals
//     lda_infer on the same document bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <numeric>
#include <thread>
#include <unordered_map>

#include "lda_host.h"

using lda_host::fail;

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Req {
    std::vector<uint32_t> words;
    std::vector<uint64_t> off;  // ndocs + 1
    std::vector<float> theta;   // ndocs x K
    std::vector<double> ll;
    int ndocs = 0, ndone = 0;
    uint64_t arena_base = 0;
    double t_submit = 0.0, t_done = 0.0;
    bool done = false;
    int status = LDA_OK;
};

struct Doc {
    uint64_t req;
    int index;   // document index inside its request
    uint64_t b;  // token range [b, b + len) in the arena
    uint32_t len;
    int segs;    // wavefronts sampling it (lda::fold_segments)
    int it_done;
};

}  // namespace

struct lda_server {
    lda_ctx* ctx = nullptr;
    int iterations = 0, step_iters = 0;
    int64_t step_tokens = 0;
    hipStream_t stream = nullptr;
    DevBuf<uint32_t> words;  // token arena
    DevBuf<uint16_t> z;
    DevBuf<lda::InferJob> jobs;
    DevBuf<lda::InferParams> params;
    DevBuf<uint32_t> counter;  // one per bucket
    lda_host::FoldStreams fs;
    uint32_t seg_tokens = 0;
    int seg_max = 1;
    DevBuf<float> theta;  // max_active x K: the finished documents of one step
    DevBuf<double> ll;
    // pinned staging
    lda::InferParams* h_params = nullptr;
    lda::InferJob* h_jobs = nullptr;
    float* h_theta = nullptr;
    double* h_ll = nullptr;
    uint64_t arena_cap = 0;
    int max_active = 0;
    int grid[lda::kFoldBuckets] = {};
    std::map<uint64_t, uint64_t> free_ranges;  // arena offset -> length
    std::deque<uint64_t> queue;                // submitted, not admitted
    std::unordered_map<uint64_t, Req> reqs;
    std::vector<Doc> active;
    uint64_t next_id = 1;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    bool stop = false;
    int fatal = LDA_OK;
    std::string fatal_msg;
    // statistics
    double busy_ms = 0.0;
    int64_t steps = 0, docs_done = 0, reqs_done = 0;
    double token_iters = 0.0;
};

namespace {

bool arena_alloc(lda_server* s, uint64_t len, uint64_t* off) {
    for (auto it = s->free_ranges.begin(); it != s->free_ranges.end(); ++it) {
        if (it->second >= len) {
            *off = it->first;
            const uint64_t rest = it->second - len, noff = it->first + len;
            s->free_ranges.erase(it);
            if (rest) s->free_ranges[noff] = rest;
            return true;
        }
    }
    return false;
}

void arena_free(lda_server* s, uint64_t off, uint64_t len) {
    if (!len) return;
    auto it = s->free_ranges.emplace(off, len).first;
    auto next = std::next(it);
    if (next != s->free_ranges.end() && it->first + it->second == next->first) {
        it->second += next->second;
        s->free_ranges.erase(next);
    }
    if (it != s->free_ranges.begin()) {
        auto prev = std::prev(it);
        if (prev->first + prev->second == it->first) {
            prev->second += it->second;
            s->free_ranges.erase(it);
        }
    }
}

// One step on the device over the staged jobs (grouped by bucket, longest
// first inside a bucket; h_params holds the bucket ranges); the first nfin
// output rows (the documents this step finishes) come back to the host.
int run_step(lda_server* s, int64_t njobs, int nfin) {
    lda_ctx* c = s->ctx;
    const hipStream_t st = s->stream;
    s->h_params->njobs = njobs;
    int grids[lda::kFoldBuckets];
    for (int b = 0; b < lda::kFoldBuckets; ++b)
        grids[b] = (int)std::min<int64_t>(s->grid[b], s->h_params->bucket_hi[b] - s->h_params->bucket_lo[b]);
    HIP_TRY(hipMemcpyAsync(s->params.p, s->h_params, sizeof(lda::InferParams), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(s->jobs.p, s->h_jobs, njobs * sizeof(lda::InferJob), hipMemcpyHostToDevice, st));
    HIP_TRY(lda::launch_fill_u32(s->counter.p, 0u, lda::kFoldBuckets, st));
    HIP_TRY(lda_host::launch_fold_buckets(c->kpl, s->params.p, grids, st, s->fs));
    HIP_TRY(lda::launch_infer_finalize(c->kpl, s->params.p, (int)std::min<int64_t>(s->grid[0], njobs), st));
    if (nfin) {
        HIP_TRY(hipMemcpyAsync(s->h_theta, s->theta.p, (size_t)nfin * c->K * sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(s->h_ll, s->ll.p, (size_t)nfin * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    return LDA_OK;
}

// Admit queued requests in FIFO order while they fit (caller holds the lock).
void admit(lda_server* s, std::vector<Doc>& newly, std::vector<Req*>& uploads) {
    while (!s->queue.empty()) {
        const uint64_t id = s->queue.front();
        Req& r = s->reqs[id];
        if ((int64_t)s->active.size() + (int64_t)newly.size() + r.ndocs > s->max_active) break;
        const uint64_t n = r.off[r.ndocs];
        uint64_t base = 0;
        if (n && !arena_alloc(s, n, &base)) break;
        r.arena_base = base;
        for (int i = 0; i < r.ndocs; ++i) {
            const uint32_t len = (uint32_t)(r.off[i + 1] - r.off[i]);
            newly.push_back(Doc{id, i, base + r.off[i], len, lda::fold_segments(len, s->seg_tokens, s->seg_max), 0});
        }
        uploads.push_back(&r);
        s->queue.pop_front();
    }
}

void fail_all(lda_server* s, int rc, double t) {
    s->fatal = rc;
    s->fatal_msg = lda_last_error();
    for (auto& kv : s->reqs) {
        if (!kv.second.done) {
            kv.second.done = true;
            kv.second.status = rc;
            kv.second.t_done = t;
        }
    }
    s->active.clear();
    s->queue.clear();
    s->cv_done.notify_all();
}

void worker_loop(lda_server* s) {
    const int K = s->ctx->K;
    if (hipSetDevice(s->ctx->device) != hipSuccess) {
        std::lock_guard<std::mutex> lk(s->mu);
        fail_all(s, fail(LDA_ERR_DEVICE, "hipSetDevice failed in the server thread"), now_ms());
    }
    for (;;) {
        std::vector<Doc> newly;
        std::vector<Req*> uploads;
        std::vector<Doc> batch;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty() || !s->active.empty(); });
            if (s->stop) return;
            admit(s, newly, uploads);
            s->active.insert(s->active.end(), newly.begin(), newly.end());
            batch = s->active;
        }
        // Token uploads: a request's documents are contiguous in the arena.
        // A Req stays in place until its waiter erases it, which needs
        // done == true, and only this thread sets that (below).
        int rc = LDA_OK;
        for (Req* r : uploads) {
            const uint64_t n = r->off[r->ndocs];
            if (n && hipMemcpyAsync(s->words.p + r->arena_base, r->words.data(), n * sizeof(uint32_t),
                                    hipMemcpyHostToDevice, s->stream) != hipSuccess) {
                rc = fail(LDA_ERR_DEVICE, "token upload failed");
                break;
            }
        }
        // This step's share of each document's iterations: at most step_tokens
        // token-iterations per wavefront.  Jobs are grouped by bucket and run
        // longest first inside a bucket; documents this step finishes take
        // output rows 0..nfin-1.
        const int64_t nb = (int64_t)batch.size();
        std::vector<int> it_end(nb);
        std::vector<int64_t> work(nb);
        int nfin = 0;
        for (int64_t i = 0; i < nb; ++i) {
            const Doc& d = batch[i];
            const int64_t per_wave = std::max<int64_t>(((int64_t)d.len + d.segs - 1) / d.segs, 1);
            int64_t chunk = s->step_tokens / per_wave;
            chunk = std::max<int64_t>(1, std::min<int64_t>(chunk, s->step_iters));
            it_end[i] = (int)std::min<int64_t>(d.it_done + chunk, s->iterations);
            work[i] = per_wave * (it_end[i] - d.it_done);
            if (it_end[i] >= s->iterations) ++nfin;
        }
        std::vector<int64_t> order(nb);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
            if (batch[a].segs != batch[b].segs) return batch[a].segs < batch[b].segs;
            return work[a] > work[b];
        });
        {
            int64_t q = 0;
            for (int bk = 0; bk < lda::kFoldBuckets; ++bk) {
                s->h_params->bucket_lo[bk] = q;
                while (q < nb && lda::fold_bucket(batch[order[q]].segs) == bk) ++q;
                s->h_params->bucket_hi[bk] = q;
            }
        }
        std::vector<int> out_row(nb, -1);
        int next_row = 0;
        for (int64_t q = 0; q < nb; ++q) {
            const int64_t i = order[q];
            const Doc& d = batch[i];
            if (it_end[i] >= s->iterations) out_row[i] = next_row++;
            s->h_jobs[q] = lda::InferJob{d.b, d.b + d.len, d.it_done, it_end[i], s->iterations,
                                         out_row[i] < 0 ? 0 : out_row[i]};
        }
        const double t0 = now_ms();
        if (rc == LDA_OK && nb) rc = run_step(s, nb, nfin);
        const double t1 = now_ms();
        std::lock_guard<std::mutex> lk(s->mu);
        s->busy_ms += t1 - t0;
        s->steps += 1;
        if (rc != LDA_OK) {
            fail_all(s, rc, t1);
            continue;
        }
        bool any = false;
        std::vector<Doc> keep;
        keep.reserve(nb);
        for (int64_t i = 0; i < nb; ++i) {
            Doc d = batch[i];
            if (out_row[i] < 0) {
                d.it_done = it_end[i];
                keep.push_back(d);
                continue;
            }
            Req& r = s->reqs[d.req];
            std::copy(s->h_theta + (size_t)out_row[i] * K, s->h_theta + (size_t)(out_row[i] + 1) * K,
                      r.theta.begin() + (size_t)d.index * K);
            r.ll[d.index] = s->h_ll[out_row[i]];
            arena_free(s, d.b, d.len);
            s->docs_done += 1;
            s->token_iters += (double)d.len * s->iterations;
            if (++r.ndone == r.ndocs) {
                r.done = true;
                r.t_done = t1;
                s->reqs_done += 1;
                any = true;
            }
        }
        s->active.swap(keep);
        if (any) s->cv_done.notify_all();
    }
}

void release(lda_server* s) {
    s->words.release(); s->z.release(); s->jobs.release(); s->params.release(); s->counter.release();
    s->theta.release(); s->ll.release();
    if (s->h_params) (void)hipHostFree(s->h_params);
    if (s->h_jobs) (void)hipHostFree(s->h_jobs);
    if (s->h_theta) (void)hipHostFree(s->h_theta);
    if (s->h_ll) (void)hipHostFree(s->h_ll);
    s->fs.destroy();
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

}  // namespace

extern "C" {

int lda_serve_start(lda_ctx* ctx, uint64_t hbm_budget_bytes, int iterations, int step_iters, int64_t step_tokens,
                    uint64_t infer_seed, lda_server** out) {
    if (!ctx || !out) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (iterations < 0) return fail(LDA_ERR_INVALID_ARGUMENT, "iterations must be >= 0");
    if (step_iters < 1 || step_tokens < 1) return fail(LDA_ERR_INVALID_ARGUMENT, "step_iters and step_tokens must be >= 1");
    if (hbm_budget_bytes < (64u << 10)) return fail(LDA_ERR_INVALID_ARGUMENT, "HBM budget below 64 KiB");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    // budget: 1/8 for per-document outputs and jobs, the rest for the token arena
    const size_t per_doc = ctx->K * sizeof(float) + sizeof(double) + sizeof(lda::InferJob);
    const size_t per_token = sizeof(uint32_t) + sizeof(uint16_t);
    lda_server* s = new lda_server();
    s->ctx = ctx;
    s->iterations = iterations;
    s->step_iters = step_iters;
    s->step_tokens = step_tokens;
    s->max_active = (int)std::min<uint64_t>(1u << 16, std::max<uint64_t>(64, hbm_budget_bytes / 8 / per_doc));
    const uint64_t doc_bytes = (uint64_t)s->max_active * per_doc;
    s->arena_cap = hbm_budget_bytes > doc_bytes ? (hbm_budget_bytes - doc_bytes) / per_token : 0;
    if (s->arena_cap < 1024) {
        release(s);
        return fail(LDA_ERR_INVALID_ARGUMENT, "HBM budget leaves no room for the token arena");
    }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess || s->fs.create() != hipSuccess) {
        release(s);
        return fail(LDA_ERR_DEVICE, "stream creation failed");
    }
    s->seg_tokens = ctx->seg_tokens;
    s->seg_max = ctx->seg_max;
    if (s->words.alloc(s->arena_cap) || s->z.alloc(s->arena_cap) || s->jobs.alloc(s->max_active) || s->params.alloc(1) ||
        s->counter.alloc(lda::kFoldBuckets) || s->theta.alloc((size_t)s->max_active * ctx->K) || s->ll.alloc(s->max_active) ||
        hipHostMalloc((void**)&s->h_params, sizeof(lda::InferParams), 0) ||
        hipHostMalloc((void**)&s->h_jobs, (size_t)s->max_active * sizeof(lda::InferJob), 0) ||
        hipHostMalloc((void**)&s->h_theta, (size_t)s->max_active * ctx->K * sizeof(float), 0) ||
        hipHostMalloc((void**)&s->h_ll, (size_t)s->max_active * sizeof(double), 0)) {
        release(s);
        return fail(LDA_ERR_OUT_OF_MEMORY, "server allocation failed");
    }
    lda::InferParams& p = *s->h_params;
    p = lda::InferParams{};
    p.words = s->words.p; p.z = s->z.p; p.jobs = s->jobs.p; p.njobs = 0;
    p.counters = s->counter.p; p.job_base = 0;
    p.nkw = ctx->d_nkw.p; p.nk = ctx->d_nk.p; p.alpha = ctx->d_alpha.p; p.alpha_cdf = ctx->d_alpha_cdf.p;
    p.eta = ctx->eta; p.veta = ctx->veta; p.K = ctx->K; p.V = ctx->V; p.Kp = ctx->Kp;
    p.seed_lo = (uint32_t)infer_seed; p.seed_hi = (uint32_t)(infer_seed >> 32);
    p.theta = s->theta.p; p.ll = s->ll.p;
    p.asum_d = 0.0; p.asum_f = 0.0f;
    for (int k = 0; k < ctx->K; ++k) {
        p.asum_d += (double)ctx->alpha[k];
        p.asum_f = p.asum_f + ctx->alpha[k];
    }
    for (int b = 0; b < lda::kFoldBuckets; ++b) s->grid[b] = ctx->fold_grid[b];
    s->free_ranges[0] = s->arena_cap;
    s->worker = std::thread(worker_loop, s);
    *out = s;
    return LDA_OK;
}

int lda_serve_submit(lda_server* s, const uint32_t* words, const uint64_t* doc_off, int64_t D, uint64_t* request_id) {
    if (!s || !request_id || D < 0 || (D && !doc_off)) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    if (D > s->max_active)
        return fail(LDA_ERR_OUT_OF_MEMORY, "request has %lld documents; the server admits at most %d at once", (long long)D,
                    s->max_active);
    if (D && doc_off[0] != 0) return fail(LDA_ERR_INVALID_ARGUMENT, "doc_off[0] must be 0");
    for (int64_t d = 0; d < D; ++d) {
        if (doc_off[d + 1] < doc_off[d]) return fail(LDA_ERR_INVALID_ARGUMENT, "doc_off must be non-decreasing");
        if (doc_off[d + 1] - doc_off[d] > lda_host::kMaxDocTokens)
            return fail(LDA_ERR_UNSUPPORTED, "document %lld has more than 2^20 tokens", (long long)d);
    }
    const uint64_t N = D ? doc_off[D] : 0;
    if (N && !words) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    for (int64_t d = 0; d < D; ++d)
        for (uint64_t g = doc_off[d] + 1; g < doc_off[d + 1]; ++g)
            if (words[g] < words[g - 1]) return fail(LDA_ERR_INVALID_ARGUMENT, "words must be sorted within each document");
    if (N > s->arena_cap)
        return fail(LDA_ERR_OUT_OF_MEMORY, "request has %llu tokens; the server's token arena holds %llu",
                    (unsigned long long)N, (unsigned long long)s->arena_cap);
    Req r;
    if (N) r.words.assign(words, words + N);
    if (D) r.off.assign(doc_off, doc_off + D + 1);
    else r.off.assign(1, 0);
    r.ndocs = (int)D;
    r.theta.assign((size_t)D * s->ctx->K, 0.0f);
    r.ll.assign(D, 0.0);
    r.t_submit = now_ms();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->fatal) return fail(s->fatal, "server stopped after an error: %s", s->fatal_msg.c_str());
        if (s->stop) return fail(LDA_ERR_STATE, "server is stopping");
        const uint64_t id = s->next_id++;
        if (!D) {
            r.done = true;
            r.t_done = r.t_submit;
        }
        s->reqs.emplace(id, std::move(r));
        if (D) s->queue.push_back(id);
        *request_id = id;
    }
    s->cv_work.notify_one();
    return LDA_OK;
}

int lda_serve_wait(lda_server* s, uint64_t request_id, int timeout_ms, float* theta, double* ll, double* latency_ms) {
    if (!s) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(request_id);
    if (it == s->reqs.end()) return fail(LDA_ERR_INVALID_ARGUMENT, "unknown request %llu", (unsigned long long)request_id);
    Req* r = &it->second;  // unordered_map nodes do not move
    auto ready = [&] { return r->done; };
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
        return fail(LDA_ERR_STATE, "request %llu has not finished", (unsigned long long)request_id);
    const int status = r->status;
    if (status == LDA_OK) {
        if (theta) std::copy(r->theta.begin(), r->theta.end(), theta);
        if (ll) std::copy(r->ll.begin(), r->ll.end(), ll);
    }
    if (latency_ms) *latency_ms = r->t_done - r->t_submit;
    s->reqs.erase(request_id);
    if (status != LDA_OK) return fail(status, "request failed: %s", s->fatal_msg.c_str());
    return LDA_OK;
}

int lda_serve_stats(lda_server* s, double stats[8]) {
    if (!s || !stats) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    stats[0] = (double)s->steps;
    stats[1] = s->busy_ms;
    stats[2] = (double)s->docs_done;
    stats[3] = (double)s->reqs_done;
    stats[4] = s->token_iters;
    stats[5] = (double)s->active.size();
    stats[6] = (double)s->queue.size();
    stats[7] = (double)s->arena_cap;
    return LDA_OK;
}

int lda_serve_stop(lda_server* s) {
    if (!s) return LDA_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->stream);
    release(s);
    return LDA_OK;
}

}  // extern "C"
