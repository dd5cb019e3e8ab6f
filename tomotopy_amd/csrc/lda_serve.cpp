// lda_serve.cpp -- continuous-batching fold-in server over a frozen model.
//
// tomotopy answers infer() calls one document set at a time (py_LDA.cpp:345-411
// -> LDAModel::_infer, LDAModel.hpp:920-953).  A server keeps the trained
// n_kw / n_k resident on the device and batches requests as they arrive:
//   * admission: requests are taken in FIFO order once their tokens fit the
//     token arena (first fit over a free list) and their documents fit the
//     active-document limit; both are sized from the HBM budget;
//   * continuous batching: a step is one persistent fold-in launch over a set
//     of admitted documents.  Steps run on kLanes lanes (streams with their
//     own buffers); whenever a lane is free and documents are waiting, the
//     waiting documents (new arrivals and those whose previous step ended)
//     start a step on it, while the other lanes' steps still run.  A
//     document advances by up to step_iters iterations per step, fewer when
//     it is long (at most step_tokens token-iterations per wavefront), so one
//     long document does not hold the others back; finished documents leave
//     after the step that ends them.  The server thread polls the lanes'
//     completion events every 100 us while steps are in flight;
//   * every draw is keyed by (infer seed, token position in the document,
//     iteration), and each launch rebuilds n_dk from the stored z, so a
//     document's result does not depend on the requests it shared the device
//     with nor on how its iterations were split into steps: it equals
//     lda_infer on the same document bit for bit.
// Term-weighted (IDF / PMI) models, K above 2048 and alpha >= 2^20 are
// served by the run-column fold-in instead: float counts depend on the order of their
// updates and cannot be rebuilt from z, so every document runs all its
// iterations within the step that starts it.  Each lane is a thread that
// takes the queued requests in FIFO order (up to the step's document and
// token limits), folds them in as one batch (lda_infer_w mode 0 on the
// lane's own buffers and streams) and hands the results back; results equal
// lda_infer_w (or lda_infer) mode 0 with the same infer_seed, bit for bit.
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
    std::vector<float> wt;      // token weights (term-weighted models)
    std::vector<uint64_t> off;  // ndocs + 1
    std::vector<float> theta;   // ndocs x K
    std::vector<double> ll;
    int ndocs = 0, ndone = 0;
    uint64_t arena_base = 0;
    bool uploaded = false;
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
    int lane;    // lane of its running step, -1 while waiting
};

constexpr int kMaxLanes = 4;

// One lane: a stream, its bucket streams, per-step buffers and the step in flight.
struct Lane {
    ::lda_server* srv = nullptr;
    hipStream_t stream = nullptr;
    lda_host::FoldStreams fs;
    hipEvent_t ev = nullptr;  // recorded after the step's results are on the host
    hipEvent_t ev0 = nullptr; // recorded before the step (timing)
    int64_t njobs = 0;
    DevBuf<lda::InferJob> jobs;
    DevBuf<lda::InferParams> params;
    DevBuf<uint32_t> counter;  // one per bucket
    DevBuf<float> theta;       // max_active x K: the documents the step finishes
    DevBuf<double> ll;
    lda::InferParams* h_params = nullptr;  // pinned staging
    lda::InferJob* h_jobs = nullptr;
    float* h_theta = nullptr;
    double* h_ll = nullptr;
    bool busy = false;
    std::vector<uint64_t> docs;  // document keys of the step
    std::vector<int> it_end, out_row;
    lda_host::GenFold gen;       // run-column steps: the lane's fold-in buffers
};

}  // namespace

struct lda_server {
    lda_ctx* ctx = nullptr;
    int iterations = 0, step_iters = 0;
    int64_t step_tokens = 0;
    uint64_t infer_seed = 0;
    DevBuf<uint32_t> words;  // token arena
    DevBuf<uint16_t> z;
    uint32_t seg_tokens = 0;
    int seg_max = 1;
    Lane lanes[kMaxLanes];
    int nlanes = 2;
    uint64_t arena_cap = 0;
    int max_active = 0;
    int grid[lda::kFoldBuckets] = {};
    std::map<uint64_t, uint64_t> free_ranges;  // arena offset -> length
    std::deque<uint64_t> queue;                // submitted, not admitted
    std::unordered_map<uint64_t, Req> reqs;
    std::map<uint64_t, Doc> active;            // admitted documents by key
    bool admit_blocked = false;                // the queue head did not fit; retry after a step ends
    uint64_t next_id = 1, next_doc = 1;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    // run-column steps (term-weighted models, K above 2048): one thread per lane
    bool generic = false, weighted = false;
    std::vector<std::thread> gworkers;
    int64_t inflight_docs = 0;
    bool stop = false;
    int fatal = LDA_OK;
    std::string fatal_msg;
    // statistics
    double busy_ms = 0.0, gpu_ms = 0.0, host_ms = 0.0, max_step_ms = 0.0;
    int64_t steps = 0, docs_done = 0, reqs_done = 0, jobs_total = 0;
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

// Launches the staged step of a lane (jobs grouped by bucket, longest first
// inside a bucket; h_params holds the bucket ranges): the first nfin output
// rows (the documents it finishes) come back to pinned host memory, then the
// lane's event is recorded.
int launch_step(lda_server* s, Lane& L, int64_t njobs, int nfin) {
    lda_ctx* c = s->ctx;
    const hipStream_t st = L.stream;
    L.h_params->njobs = njobs;
    int grids[lda::kFoldBuckets];
    for (int b = 0; b < lda::kFoldBuckets; ++b)
        grids[b] = (int)std::min<int64_t>(s->grid[b], L.h_params->bucket_hi[b] - L.h_params->bucket_lo[b]);
    HIP_TRY(hipEventRecord(L.ev0, st));
    HIP_TRY(hipMemcpyAsync(L.params.p, L.h_params, sizeof(lda::InferParams), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(L.jobs.p, L.h_jobs, njobs * sizeof(lda::InferJob), hipMemcpyHostToDevice, st));
    HIP_TRY(lda::launch_fill_u32(L.counter.p, 0u, lda::kFoldBuckets, st));
    HIP_TRY(lda_host::launch_fold_buckets(c->kpl, c->has_priors, L.params.p, grids, st, L.fs));
    HIP_TRY(lda::launch_infer_finalize(c->kpl, c->has_priors, L.params.p, (int)std::min<int64_t>(s->grid[0], njobs), st));
    if (nfin) {
        HIP_TRY(hipMemcpyAsync(L.h_theta, L.theta.p, (size_t)nfin * c->K * sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(L.h_ll, L.ll.p, (size_t)nfin * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipEventRecord(L.ev, st));
    return LDA_OK;
}

// Admit queued requests in FIFO order while they fit (caller holds the lock).
void admit(lda_server* s) {
    while (!s->queue.empty()) {
        const uint64_t id = s->queue.front();
        Req& r = s->reqs[id];
        const uint64_t n = r.off[r.ndocs];
        uint64_t base = 0;
        if ((int64_t)s->active.size() + r.ndocs > s->max_active || (n && !arena_alloc(s, n, &base))) {
            s->admit_blocked = true;
            return;
        }
        r.arena_base = base;
        for (int i = 0; i < r.ndocs; ++i) {
            const uint32_t len = (uint32_t)(r.off[i + 1] - r.off[i]);
            s->active.emplace(s->next_doc++,
                              Doc{id, i, base + r.off[i], len, lda::fold_segments(len, s->seg_tokens, s->seg_max), 0, -1});
        }
        s->queue.pop_front();
    }
    s->admit_blocked = false;
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

// Results of a finished lane: finished documents go to their requests, the
// others wait for their next step (caller holds the lock).
void finish_lane(lda_server* s, Lane& L, double t) {
    const int K = s->ctx->K;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, L.ev0, L.ev) == hipSuccess) {
        s->gpu_ms += ms;
        s->max_step_ms = std::max(s->max_step_ms, (double)ms);
    }
    bool any = false;
    s->admit_blocked = false;  // memory and document slots may have been freed
    for (size_t i = 0; i < L.docs.size(); ++i) {
        auto it = s->active.find(L.docs[i]);
        if (it == s->active.end()) continue;  // dropped after an error
        Doc& d = it->second;
        d.lane = -1;
        if (L.out_row[i] < 0) {
            d.it_done = L.it_end[i];
            continue;
        }
        Req& r = s->reqs[d.req];
        const int row = L.out_row[i];
        std::copy(L.h_theta + (size_t)row * K, L.h_theta + (size_t)(row + 1) * K, r.theta.begin() + (size_t)d.index * K);
        r.ll[d.index] = L.h_ll[row];
        arena_free(s, d.b, d.len);
        s->docs_done += 1;
        s->token_iters += (double)d.len * s->iterations;
        if (++r.ndone == r.ndocs) {
            r.done = true;
            r.t_done = t;
            s->reqs_done += 1;
            any = true;
        }
        s->active.erase(it);
    }
    L.docs.clear();
    L.busy = false;
    if (any) s->cv_done.notify_all();
}

// Stages a step of every waiting document on lane L (caller holds the lock):
// at most step_tokens token-iterations per wavefront and step_iters
// iterations per document; jobs grouped by bucket and longest first inside a
// bucket; documents the step finishes take output rows 0..nfin-1.  Returns
// the number of jobs; *uploads receives the requests whose tokens are not on
// the device yet.
int64_t stage_step(lda_server* s, Lane& L, int* nfin_out, std::vector<Req*>* uploads) {
    std::vector<uint64_t> keys;
    for (auto& kv : s->active)
        if (kv.second.lane < 0) keys.push_back(kv.first);
    const int64_t nb = (int64_t)keys.size();
    std::vector<const Doc*> dd(nb);
    std::vector<int> it_end(nb);
    std::vector<int64_t> work(nb);
    int nfin = 0;
    for (int64_t i = 0; i < nb; ++i) {
        const Doc& d = s->active[keys[i]];
        dd[i] = &d;
        const int64_t per_wave = std::max<int64_t>(((int64_t)d.len + d.segs - 1) / d.segs, 1);
        int64_t chunk = s->step_tokens / per_wave;
        chunk = std::max<int64_t>(1, std::min<int64_t>(chunk, s->step_iters));
        it_end[i] = (int)std::min<int64_t>(d.it_done + chunk, s->iterations);
        work[i] = per_wave * (it_end[i] - d.it_done);
        if (it_end[i] >= s->iterations) ++nfin;
        Req& r = s->reqs[d.req];
        if (!r.uploaded) {
            r.uploaded = true;
            uploads->push_back(&r);
        }
    }
    std::vector<int64_t> order(nb);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        if (dd[a]->segs != dd[b]->segs) return dd[a]->segs < dd[b]->segs;
        return work[a] > work[b];
    });
    int64_t q = 0;
    for (int bk = 0; bk < lda::kFoldBuckets; ++bk) {
        L.h_params->bucket_lo[bk] = q;
        while (q < nb && lda::fold_bucket(dd[order[q]]->segs) == bk) ++q;
        L.h_params->bucket_hi[bk] = q;
    }
    L.docs.resize(nb);
    L.it_end.resize(nb);
    L.out_row.resize(nb);
    int next_row = 0;
    for (int64_t j = 0; j < nb; ++j) {
        const int64_t i = order[j];
        const Doc& d = *dd[i];
        const bool fin = it_end[i] >= s->iterations;
        L.docs[j] = keys[i];
        L.it_end[j] = it_end[i];
        L.out_row[j] = fin ? next_row++ : -1;
        L.h_jobs[j] = lda::InferJob{d.b, d.b + d.len, d.it_done, it_end[i], s->iterations, fin ? L.out_row[j] : 0};
        s->active[keys[i]].lane = (int)(&L - s->lanes);
    }
    *nfin_out = nfin;
    return nb;
}

void worker_loop(lda_server* s) {
    if (hipSetDevice(s->ctx->device) != hipSuccess) {
        std::lock_guard<std::mutex> lk(s->mu);
        fail_all(s, fail(LDA_ERR_DEVICE, "hipSetDevice failed in the server thread"), now_ms());
    }
    std::unique_lock<std::mutex> lk(s->mu);
    auto waiting = [&] {
        for (auto& kv : s->active)
            if (kv.second.lane < 0) return true;
        return false;
    };
    auto free_lane = [&]() -> Lane* {
        for (int i = 0; i < s->nlanes; ++i)
            if (!s->lanes[i].busy) return &s->lanes[i];
        return nullptr;
    };
    auto in_flight = [&] {
        for (int i = 0; i < s->nlanes; ++i)
            if (s->lanes[i].busy) return true;
        return false;
    };
    while (!s->stop) {
        const double t = now_ms();
        for (int i = 0; i < s->nlanes; ++i) {
            Lane& L = s->lanes[i];
            if (!L.busy) continue;
            const hipError_t q = hipEventQuery(L.ev);
            if (q == hipSuccess) finish_lane(s, L, t);
            else if (q != hipErrorNotReady) fail_all(s, fail(LDA_ERR_DEVICE, "step failed: %s", hipGetErrorString(q)), t);
        }
        if (!s->fatal) admit(s);
        Lane* L = s->fatal ? nullptr : free_lane();
        if (!L || !waiting()) {
            if (in_flight()) s->cv_work.wait_for(lk, std::chrono::microseconds(100));
            else s->cv_work.wait(lk, [&] { return s->stop || (!s->queue.empty() && !s->admit_blocked); });
            continue;
        }
        const double th0 = now_ms();
        int nfin = 0;
        std::vector<Req*> uploads;
        const int64_t nb = stage_step(s, *L, &nfin, &uploads);
        L->busy = true;
        L->njobs = nb;
        s->steps += 1;
        s->jobs_total += nb;
        // Requests stay in place until their waiter erases them, which needs
        // done == true, set only by this thread: the uploads can run unlocked.
        lk.unlock();
        int rc = LDA_OK;
        for (Req* r : uploads) {
            const uint64_t n = r->off[r->ndocs];
            if (n && hipMemcpyAsync(s->words.p + r->arena_base, r->words.data(), n * sizeof(uint32_t),
                                    hipMemcpyHostToDevice, L->stream) != hipSuccess) {
                rc = fail(LDA_ERR_DEVICE, "token upload failed");
                break;
            }
        }
        if (rc == LDA_OK) rc = launch_step(s, *L, nb, nfin);
        lk.lock();
        s->host_ms += now_ms() - th0;
        if (rc != LDA_OK) {
            L->busy = false;
            fail_all(s, rc, now_ms());
        }
    }
}

// A lane of the run-column server: takes queued requests (FIFO, at most
// max_active documents and arena_cap tokens per step, at least one request),
// folds them in as one batch and returns the results.
void worker_loop_gen(lda_server* s, int lane) {
    Lane& L = s->lanes[lane];
    lda_ctx* c = s->ctx;
    if (hipSetDevice(c->device) != hipSuccess) {
        std::lock_guard<std::mutex> lk(s->mu);
        fail_all(s, fail(LDA_ERR_DEVICE, "hipSetDevice failed in the server thread"), now_ms());
        return;
    }
    std::vector<uint32_t> words;
    std::vector<float> wts;
    std::vector<uint64_t> off;
    std::vector<float> theta;
    std::vector<double> ll;
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_work.wait(lk, [&] { return s->stop || (!s->queue.empty() && !s->fatal); });
        if (s->stop) break;
        const double th0 = now_ms();
        std::vector<Req*> batch;
        int64_t nd = 0;
        uint64_t nt = 0;
        while (!s->queue.empty()) {
            Req& r = s->reqs[s->queue.front()];
            const uint64_t n = r.off[r.ndocs];
            if (!batch.empty() && (nd + r.ndocs > s->max_active || nt + n > s->arena_cap)) break;
            batch.push_back(&r);
            nd += r.ndocs;
            nt += n;
            s->queue.pop_front();
        }
        s->inflight_docs += nd;
        lk.unlock();
        // Requests stay in place until their waiter erases them, which needs
        // done == true, set only under the lock below.
        words.resize(nt);
        wts.resize(s->weighted ? nt : 0);
        off.assign(1, 0);
        for (Req* r : batch) {
            const uint64_t base = off.back();
            const uint64_t n = r->off[r->ndocs];
            std::copy(r->words.begin(), r->words.end(), words.begin() + base);
            if (s->weighted) std::copy(r->wt.begin(), r->wt.end(), wts.begin() + base);
            for (int i = 1; i <= r->ndocs; ++i) off.push_back(base + r->off[i]);
            (void)n;
        }
        theta.resize((size_t)nd * c->K);
        ll.resize(nd);
        float ms = 0.0f;
        const int rc = lda_host::infer_general_on(c, L.gen, L.stream, L.fs, L.ev0, L.ev, words.data(), off.data(), nd,
                                                  s->weighted ? wts.data() : nullptr, s->iterations, s->infer_seed, 0,
                                                  theta.data(), ll.data(), nullptr, &ms);
        const double t = now_ms();
        lk.lock();
        s->inflight_docs -= nd;
        if (rc != LDA_OK) {
            fail_all(s, rc, t);
            continue;
        }
        int64_t row = 0;
        for (Req* r : batch) {
            std::copy(theta.begin() + row * c->K, theta.begin() + (row + r->ndocs) * c->K, r->theta.begin());
            std::copy(ll.begin() + row, ll.begin() + row + r->ndocs, r->ll.begin());
            row += r->ndocs;
            r->ndone = r->ndocs;
            r->done = true;
            r->t_done = t;
        }
        s->steps += 1;
        s->jobs_total += nd;
        s->docs_done += nd;
        s->reqs_done += (int64_t)batch.size();
        s->token_iters += (double)nt * s->iterations;
        s->gpu_ms += ms;
        s->max_step_ms = std::max(s->max_step_ms, (double)ms);
        s->busy_ms += t - th0;
        s->host_ms += std::max(0.0, t - th0 - (double)ms);
        s->cv_done.notify_all();
    }
}

void release(lda_server* s) {
    s->words.release();
    s->z.release();
    for (Lane& L : s->lanes) {
        L.jobs.release(); L.params.release(); L.counter.release(); L.theta.release(); L.ll.release();
        L.gen.release();
        if (L.h_params) (void)hipHostFree(L.h_params);
        if (L.h_jobs) (void)hipHostFree(L.h_jobs);
        if (L.h_theta) (void)hipHostFree(L.h_theta);
        if (L.h_ll) (void)hipHostFree(L.h_ll);
        L.fs.destroy();
        if (L.ev) (void)hipEventDestroy(L.ev);
        if (L.ev0) (void)hipEventDestroy(L.ev0);
        if (L.stream) (void)hipStreamDestroy(L.stream);
    }
    delete s;
}

}  // namespace

extern "C" {

int lda_serve_start(lda_ctx* ctx, uint64_t hbm_budget_bytes, int iterations, int step_iters, int64_t step_tokens,
                    int lanes, uint64_t infer_seed, lda_server** out) {
    if (!ctx || !out) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (iterations < 0) return fail(LDA_ERR_INVALID_ARGUMENT, "iterations must be >= 0");
    if (step_iters < 1 || step_tokens < 1) return fail(LDA_ERR_INVALID_ARGUMENT, "step_iters and step_tokens must be >= 1");
    if (lanes < 1 || lanes > kMaxLanes) return fail(LDA_ERR_INVALID_ARGUMENT, "lanes must be in 1..%d", kMaxLanes);
    if (hbm_budget_bytes < (64u << 10)) return fail(LDA_ERR_INVALID_ARGUMENT, "HBM budget below 64 KiB");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->tw != LDA_TW_ONE || ctx->kpl > lda::kSmallMaxKpl || ctx->exact_alpha) {
        // run-column steps: per lane and step at most max_active documents
        // and arena_cap tokens, sized so that a step's private columns (at
        // most one Kp column per token), state rows and arrays fit the budget
        lda_server* s = new lda_server();
        s->ctx = ctx;
        s->generic = true;
        s->weighted = ctx->tw != LDA_TW_ONE;
        s->iterations = iterations;
        s->step_iters = step_iters;
        s->step_tokens = step_tokens;
        s->infer_seed = infer_seed;
        s->nlanes = lanes;
        const uint64_t lane_budget = hbm_budget_bytes / (uint64_t)lanes;
        const uint64_t per_doc = (uint64_t)ctx->Kp * 8 + (uint64_t)ctx->K * 4 + 256;
        const uint64_t per_token = (uint64_t)ctx->Kp * 4 + 24;
        s->max_active = (int)std::min<uint64_t>(1u << 16, std::max<uint64_t>(1, lane_budget / 4 / per_doc));
        s->arena_cap = std::max<uint64_t>(1, (lane_budget - std::min(lane_budget, (uint64_t)s->max_active * per_doc)) / per_token);
        for (int i = 0; i < s->nlanes; ++i) {
            Lane& L = s->lanes[i];
            L.srv = s;
            if (hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) != hipSuccess || L.fs.create() != hipSuccess ||
                hipEventCreate(&L.ev) != hipSuccess || hipEventCreate(&L.ev0) != hipSuccess) {
                release(s);
                return fail(LDA_ERR_DEVICE, "stream creation failed");
            }
        }
        for (int i = 0; i < s->nlanes; ++i) s->gworkers.emplace_back(worker_loop_gen, s, i);
        *out = s;
        return LDA_OK;
    }
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
    s->seg_tokens = ctx->seg_tokens;
    s->seg_max = ctx->seg_max;
    s->infer_seed = infer_seed;
    s->nlanes = lanes;
    if (s->words.alloc(s->arena_cap) || s->z.alloc(s->arena_cap)) {
        release(s);
        return fail(LDA_ERR_OUT_OF_MEMORY, "server allocation failed");
    }
    double asd = 0.0;
    float asf = 0.0f;
    for (int k = 0; k < ctx->K; ++k) {
        asd += (double)ctx->alpha[k];
        asf = asf + ctx->alpha[k];
    }
    for (int i = 0; i < s->nlanes; ++i) {
        Lane& L = s->lanes[i];
        L.srv = s;
        if (hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) != hipSuccess || L.fs.create() != hipSuccess ||
            hipEventCreate(&L.ev) != hipSuccess || hipEventCreate(&L.ev0) != hipSuccess) {
            release(s);
            return fail(LDA_ERR_DEVICE, "stream creation failed");
        }
        if (L.jobs.alloc(s->max_active) || L.params.alloc(1) || L.counter.alloc(lda::kFoldBuckets) ||
            L.theta.alloc((size_t)s->max_active * ctx->K) || L.ll.alloc(s->max_active) ||
            hipHostMalloc((void**)&L.h_params, sizeof(lda::InferParams), 0) ||
            hipHostMalloc((void**)&L.h_jobs, (size_t)s->max_active * sizeof(lda::InferJob), 0) ||
            hipHostMalloc((void**)&L.h_theta, (size_t)s->max_active * ctx->K * sizeof(float), 0) ||
            hipHostMalloc((void**)&L.h_ll, (size_t)s->max_active * sizeof(double), 0)) {
            release(s);
            return fail(LDA_ERR_OUT_OF_MEMORY, "server allocation failed");
        }
        lda::InferParams& p = *L.h_params;
        p = lda::InferParams{};
        p.words = s->words.p; p.z = s->z.p; p.jobs = L.jobs.p; p.njobs = 0;
        p.counters = L.counter.p; p.job_base = 0;
        p.nkw = ctx->d_nkw.p; p.nk = ctx->d_nk.p; p.alpha = ctx->d_alpha.p; p.alpha_cdf = ctx->d_alpha_cdf.p;
        p.eta = ctx->eta; p.veta = ctx->veta; p.K = ctx->K; p.V = ctx->V; p.Kp = ctx->Kp;
        p.seed_lo = (uint32_t)infer_seed; p.seed_hi = (uint32_t)(infer_seed >> 32);
        p.theta = L.theta.p; p.ll = L.ll.p;
        p.asum_d = asd; p.asum_f = asf;
        if (ctx->has_priors) {
            p.prior_row = ctx->d_prior_row.p; p.prior_eta = ctx->d_prior_eta.p;
            p.etasum = ctx->d_etasum.p; p.etasum_d = ctx->d_etasum_d.p;
        }
    }
    for (int b = 0; b < lda::kFoldBuckets; ++b) s->grid[b] = ctx->fold_grid[b];
    s->free_ranges[0] = s->arena_cap;
    s->worker = std::thread(worker_loop, s);
    *out = s;
    return LDA_OK;
}

}  // extern "C"

namespace {

int submit_impl(lda_server* s, const uint32_t* words, const uint64_t* doc_off, int64_t D, const float* weights,
                uint64_t* request_id) {
    if (!s || !request_id || D < 0 || (D && !doc_off)) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    if (s->weighted && D && doc_off[D] && !weights)
        return fail(LDA_ERR_INVALID_ARGUMENT, "a term-weighted model's server needs the token weights (lda_serve_submit_w)");
    if (!s->weighted && weights) return fail(LDA_ERR_INVALID_ARGUMENT, "TermWeight ONE models take no token weights");
    if (D > s->max_active)
        return fail(LDA_ERR_OUT_OF_MEMORY, "request has %lld documents; the server admits at most %d at once", (long long)D,
                    s->max_active);
    if (D && doc_off[0] != 0) return fail(LDA_ERR_INVALID_ARGUMENT, "doc_off[0] must be 0");
    for (int64_t d = 0; d < D; ++d) {
        if (doc_off[d + 1] < doc_off[d]) return fail(LDA_ERR_INVALID_ARGUMENT, "doc_off must be non-decreasing");
        if (!s->generic && doc_off[d + 1] - doc_off[d] > lda_host::kMaxDocTokens)
            return fail(LDA_ERR_UNSUPPORTED,
                        "document %lld has more than 2^20 tokens: this server keeps float(n_dk) + alpha per topic; "
                        "fold it in with lda_infer", (long long)d);
    }
    const uint64_t N = D ? doc_off[D] : 0;
    if (N && !words) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    for (int64_t d = 0; d < D; ++d)
        for (uint64_t g = doc_off[d] + 1; g < doc_off[d + 1]; ++g)
            if (words[g] < words[g - 1]) return fail(LDA_ERR_INVALID_ARGUMENT, "words must be sorted within each document");
    if (N > s->arena_cap)
        return fail(LDA_ERR_OUT_OF_MEMORY, "request has %llu tokens; the server's token arena holds %llu",
                    (unsigned long long)N, (unsigned long long)s->arena_cap);
    if (weights)
        for (uint64_t g = 0; g < N; ++g)
            if (words[g] < (uint32_t)s->ctx->V && !(weights[g] >= 0.0f && weights[g] < 1e30f))
                return fail(LDA_ERR_INVALID_ARGUMENT, "token weight %llu is negative or not finite", (unsigned long long)g);
    Req r;
    if (N) r.words.assign(words, words + N);
    if (N && weights) r.wt.assign(weights, weights + N);
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

}  // namespace

extern "C" {

int lda_serve_submit(lda_server* s, const uint32_t* words, const uint64_t* doc_off, int64_t D, uint64_t* request_id) {
    return submit_impl(s, words, doc_off, D, nullptr, request_id);
}

int lda_serve_submit_w(lda_server* s, const uint32_t* words, const uint64_t* doc_off, int64_t D, const float* weights,
                       uint64_t* request_id) {
    return submit_impl(s, words, doc_off, D, weights, request_id);
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

int lda_serve_stats(lda_server* s, double stats[12]) {
    if (!s || !stats) return fail(LDA_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    stats[0] = (double)s->steps;
    stats[1] = s->busy_ms;
    stats[2] = (double)s->docs_done;
    stats[3] = (double)s->reqs_done;
    stats[4] = s->token_iters;
    stats[5] = s->generic ? (double)s->inflight_docs : (double)s->active.size();
    stats[6] = (double)s->queue.size();
    stats[7] = (double)s->arena_cap;
    stats[8] = s->gpu_ms;
    stats[9] = s->host_ms;
    stats[10] = s->max_step_ms;
    stats[11] = s->steps ? (double)s->jobs_total / (double)s->steps : 0.0;
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
    for (auto& t : s->gworkers)
        if (t.joinable()) t.join();
    (void)hipSetDevice(s->ctx->device);
    for (Lane& L : s->lanes)
        if (L.stream) (void)hipStreamSynchronize(L.stream);
    release(s);
    return LDA_OK;
}

}  // extern "C"
