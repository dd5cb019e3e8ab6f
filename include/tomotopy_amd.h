/*
 * tomotopy_amd.h -- C ABI of the MI355X collapsed-Gibbs LDA engine.
 *
 * Plain C: pointers, sizes and integer status codes; no C++ exceptions and no
 * torch types cross this boundary.  Every call returns LDA_OK (0) or a
 * negative LDA_ERR_* code; lda_last_error() gives a thread-local message.
 * One context owns one model's device state and one HIP stream; calls on one
 * context must be serialised by the caller (the Python LDAModel does this).
 *
 * Layout conventions at this boundary (the device layout is private):
 *   words      uint32 token word ids, documents concatenated (CSR), each
 *              document's words sorted ascending; ids >= V are kept and
 *              skipped by the sampler (tomoto's "vid >= realV").
 *   doc_off    uint64[D+1] CSR offsets into words.
 *   z          uint16 per token; 0xFFFF for skipped tokens.
 *   nkw        int32 V x K, word-major (tomoto's column-major K x V
 *              numByTopicWord, LDAModel.hpp:66).
 *   nk         int32 K.   ndk  int32 D x K (document-major).
 *   alpha      float K.
 *
 * Reference interfaces replaced (bab2min/tomotopy):
 *   lda_ctx_create        ILDAModel::create + LDAModel ctor checks (TermWeight ONE/IDF/PMI)
 *                         (src/TopicModel/LDAModel.cpp:10-13, LDAModel.hpp:964-984)
 *   lda_upload_corpus     prepare(): initializeDocState / addWordTo<+1>
 *                         (LDAModel.hpp:716-813, 1073-1138)
 *   lda_train             ITopicModel::train (TopicModel.hpp:684-755) ->
 *                         trainOne (LDAModel.hpp:505-539) -> performSampling
 *                         (356-428) -> sampleDocument (326-354) ->
 *                         getZLikelihoods (273-285) -> sample::prefixSum /
 *                         sampleFromDiscreteAcc (src/Utils/sample.hpp:61-189)
 *   lda_sweep             one trainOne without the alpha step
 *   lda_optimize_alpha    optimizeParameters (LDAModel.hpp:221-265)
 *   lda_loglik            getLLDocs / getLLRest (LDAModel.hpp:616-674), FP64
 *   lda_infer             LDAModel::_infer<together=false> (LDAModel.hpp:843-954)
 *                         + _getTopicsByDoc (1145-1159)
 *   lda_download_state /  globalState / docs read-out for save, get_topic_words,
 *   lda_set_state         get_count_by_topics (TopicModel.hpp:335-395,
 *                         LDAModel.hpp:815-841) and load (prepare(false),
 *                         LDAModel.hpp:1064-1071, 1163-1173)
 *   lda_sample_frozen     one getZLikelihoods + sampleFromDiscreteAcc step on a
 *                         frozen state (the pure per-token function)
 */
#ifndef TOMOTOPY_AMD_H
#define TOMOTOPY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lda_ctx lda_ctx;

enum {
    LDA_OK = 0,
    LDA_ERR_INVALID_ARGUMENT = -1, /* tomoto exc::InvalidArgument -> ValueError */
    LDA_ERR_STATE = -2,            /* call order (e.g. train before upload) -> RuntimeError */
    LDA_ERR_DEVICE = -3,           /* HIP runtime error */
    LDA_ERR_OUT_OF_MEMORY = -4,
    LDA_ERR_UNSUPPORTED = -5,
    LDA_ERR_TRAINING = -6          /* train failed (tomoto returns <0 -> "Train failed") */
};

/* tomotopy.TermWeight */
enum { LDA_TW_ONE = 0, LDA_TW_IDF = 1, LDA_TW_PMI = 2 };

/* tomotopy.ParallelScheme (tomotopy/__init__.py:36-64) */
enum { LDA_PS_DEFAULT = 0, LDA_PS_NONE = 1, LDA_PS_COPY_MERGE = 2, LDA_PS_PARTITION = 3 };

/* Largest K handled by the register-resident sampler (64 lanes x 32 topics). */
#define LDA_MAX_K 32767

const char* lda_last_error(void);
int lda_version(void);

/* Number of visible HIP devices (0 when none). */
int lda_device_count(int* count);

/* Create a model context on `device`.  alpha: K floats (symmetric alpha is K
 * copies).  eta > 0, 1 <= K <= LDA_MAX_K, V >= 0.  termweight: LDA_TW_*. */
int lda_ctx_create(int K, int V, const float* alpha, float eta, int termweight, uint64_t seed, int device,
                   lda_ctx** out);
int lda_ctx_destroy(lda_ctx* ctx);

/* A deep copy of the context on the same device: configuration, corpus,
 * topics, every count buffer (integer or float), word priors, alpha, the
 * global step and the fold-in options, copied device to device, so that the
 * copy trains and folds in exactly as the original would.  Replaces
 * TopicModel::copy (TopicModel.hpp:623-628: the derived model's copy
 * constructor; py LDAModel.copy, tomotopy/models.py:295-299). */
int lda_ctx_copy(lda_ctx* ctx, lda_ctx** out);

/* Test hook (no counterpart in the reference; for tests of the error path
 * that tomotopy maps to RuntimeError("Train failed"), py_LDA.cpp:306-309):
 * the lda_train iteration that starts at global step `step` fails with
 * LDA_ERR_TRAINING.  step < 0 clears it.  Applies to this context only. */
int lda_debug_fail_at_step(lda_ctx* ctx, int64_t step);

/* Upload the training corpus.  If z is NULL the initial topics are drawn on
 * the device (z ~ Discrete(alpha) from the Philox stream); otherwise z is
 * taken as given.  Counts n_dk, n_kw, n_k are then built from z.
 * weights: N floats, required for LDA_TW_IDF / LDA_TW_PMI models (tomoto's
 * per-token wordWeights, LDAModel.hpp:802-809, 1104-1121; entries of words
 * >= V are ignored), NULL for LDA_TW_ONE.  Term-weighted models keep float
 * counts (addWordTo adds the weight; removal is max(count - weight, 0),
 * Utils.hpp:132-155) and support training under every scheme, the
 * log-likelihood and the alpha fixed point; their fold-in, frozen sampling
 * and exact invariant checks return LDA_ERR_UNSUPPORTED. */
int lda_upload_corpus(lda_ctx* ctx, const uint32_t* words, const uint64_t* doc_off, int64_t D, const uint16_t* z,
                      const float* weights);

/* Replace the topic-word counts (tomoto load takes n_kw / n_k from the file:
 * TopicModel.hpp:361-395, globalState read by ModelStateLDA's serializer,
 * LDAModel.hpp:58-69).  nkw: V x K word-major, nk: K. */
int lda_set_counts(lda_ctx* ctx, const int32_t* nkw, const int32_t* nk);
/* The same for term-weighted models (float counts). */
int lda_set_counts_f(lda_ctx* ctx, const float* nkw, const float* nk);

/* Word priors (LDAModel::setWordPrior / prepareWordPriors, LDAModel.hpp:724-737,
 * 1031-1045): word_ids[i] (ids >= V are skipped, as tomoto skips removed
 * words) gets the topic prior priors[i*K .. i*K+K) (>= 0) in place of eta;
 * every other word keeps eta.  The weights then use eta_kv and the per-topic
 * sum of eta_kv (EtaHelper<_, true>, LDAModel.hpp:99-117), the initial topics
 * are drawn from the word's eta column (LDAModel.hpp:769-773) and the
 * log-likelihood uses eta_kv (LDAModel.hpp:638-650).  Call before
 * lda_upload_corpus; n == 0 clears the priors.  Training, fold-in (alone and
 * together) and serving all use eta_kv. */
int lda_set_word_priors(lda_ctx* ctx, int64_t n, const uint32_t* word_ids, const float* priors);

int lda_set_alpha(lda_ctx* ctx, const float* alpha);
int lda_get_alpha(lda_ctx* ctx, float* alpha);
int lda_set_optim(lda_ctx* ctx, uint32_t optim_interval, uint32_t burn_in);
int lda_get_global_step(lda_ctx* ctx, uint64_t* step);
int lda_set_global_step(lda_ctx* ctx, uint64_t step);

/* Training.  scheme is an LDA_PS_* value; workers as in tomotopy:
 *   NONE, or workers == 1      one device worker, sequential over shuffled
 *                              documents (deterministic).
 *   PARTITION / DEFAULT with workers in [2, 1024]
 *                              the PCGS partition schedule with `workers`
 *                              device workers (capped at max((V+3)/4, 1) as
 *                              tomoto caps it), W rounds per sweep replayed
 *                              from a graph, private n_k per worker
 *                              (deterministic).
 *   PARTITION with workers == 0
 *                              the same schedule with the device's worker
 *                              count: the largest W that leaves each round
 *                              about 48 or more tokens per worker, within
 *                              tomoto's cap, the document count and the
 *                              resident capacity (lda_resolve_schedule
 *                              reports it).
 *   COPY_MERGE with workers = W in [2, 64]
 *                              tomoto's copy_merge: W device workers, each
 *                              with private copies of n_kw and n_k, merged
 *                              by summing deltas (deterministic); workers ==
 *                              0 picks W = min(64, (D+1)/2, copies in 32 GB).
 *   DEFAULT with workers == 0  device-wide (no counterpart in tomotopy,
 *                              whose DEFAULT is PARTITION): every workgroup
 *                              is a worker with a private n_k; n_kw is
 *                              shared and updated with exact int32 atomics,
 *                              each wavefront sending those of a group of
 *                              tokens at once (lda_set_atomic_group);
 *                              the column of the current word is read when
 *                              the word's run starts; at most ceil(D / 8)
 *                              documents in flight unless
 *                              lda_set_device_schedule says otherwise.
 * lda_train runs `iterations` sweeps, each followed by the alpha fixed point
 * when global_step >= burn_in and (global_step+1) % optim_interval == 0. */
int lda_train(lda_ctx* ctx, int iterations, int workers, int scheme);
/* The schedule lda_train(workers, scheme) runs on this context: *scheme_out
 * LDA_PS_NONE, LDA_PS_PARTITION or LDA_PS_COPY_MERGE with *workers_out
 * workers, or LDA_PS_DEFAULT for the device-wide schedule with *workers_out
 * workgroups.  Mirrors TopicModel::train's resolution (TopicModel.hpp:684-690,
 * getRealScheme 664-682, estimateMaxThreads LDAModel.hpp:491-503). */
int lda_resolve_schedule(lda_ctx* ctx, int workers, int scheme, int* scheme_out, int* workers_out);
/* Concurrency of the device-wide schedule (no counterpart in the reference):
 * at most max_workers workgroups (0: as many as are resident), each of
 * waves_per_worker (1..8) wavefronts that share one private n_k.
 * max_workers < 0 restores the default, which keeps at most ceil(D / 8)
 * documents in flight (documents sampled at the same time miss each other's
 * count changes; an eighth of the corpus keeps ll_per_word within about
 * 0.15% of tomoto's partition chain) and runs the full grid of 8-wavefront
 * workgroups once D is at least 8x the resident wavefronts. */
int lda_set_device_schedule(lda_ctx* ctx, int max_workers, int waves_per_worker);
int lda_sweep(lda_ctx* ctx, int workers, int scheme);
int lda_optimize_alpha(lda_ctx* ctx);

/* FP64 log-likelihood: ll_docs = getLLDocs(all docs), ll_rest = getLLRest. */
int lda_loglik(lda_ctx* ctx, double* ll_docs, double* ll_rest);

/* Count invariants, counted on the device.  flags[0] docs whose sum_k n_dk !=
 * valid length; [1] topics whose sum_v n_kw != n_k; [2] sum_k n_k - N_valid;
 * [3] n_dk cells differing from a recount from z; [4] n_kw cells differing
 * from a recount from z; [5] negative cells; [6] tokens with invalid z. */
int lda_check_invariants(lda_ctx* ctx, int64_t flags[8]);

/* Read-out (host buffers; any pointer may be NULL). */
int lda_download_state(lda_ctx* ctx, uint16_t* z, int32_t* ndk, int32_t* nkw, int32_t* nk);
int lda_get_dims(lda_ctx* ctx, int64_t dims[6]); /* K, V, D, N, Kp, kpl */
/* Read-out of a term-weighted model (float counts, natural topic order). */
int lda_download_state_f(lda_ctx* ctx, uint16_t* z, float* ndk, float* nkw, float* nk);
/* tomoto's weightedN (TopicModel.hpp:507-524): the token count for
 * LDA_TW_ONE, the sum of the valid tokens' weights otherwise; ll_per_word is
 * (ll_docs + ll_rest) / weightedN. */
int lda_get_weighted_n(lda_ctx* ctx, double* weighted_n);

/* Self-test of the device's reciprocal (no counterpart in the reference): the
 * sampler takes RN(1/(n_k + V*eta)) as the hardware reciprocal refined by one
 * Newton step; this counts, over all 2^23 mantissas of the binade
 * [2^exp_shift, 2^(exp_shift+1)), where it differs from IEEE division. */
int lda_selftest_reciprocal(int exp_shift, int device, uint64_t* mismatches);

/* Read-outs on the device (the whole n_kw / n_dk / z stay in HBM):
 *   lda_topic_word_dist  topic's (n_kw + eta) / (n_k + V * eta) for every word
 *                        (out: V floats; normalize = 0: n_kw + eta), in
 *                        tomoto's float operations (_getWidsByTopic,
 *                        LDAModel.hpp:828-840; TopicModel::getWidsByTopic,
 *                        TopicModel.hpp:777-780).
 *   lda_count_by_topic   tokens per topic over the words < V (out: K; tomoto's
 *                        _getTopicsCount, LDAModel.hpp:815-826).
 *   lda_doc_topic_dist   training document doc's (n_dk + alpha) / (weight sum
 *                        + sum alpha) (out: K; normalize = 0: n_dk + alpha;
 *                        _getTopicsByDoc, LDAModel.hpp:1145-1159); the weight
 *                        sum is the document's valid token count, or for IDF /
 *                        PMI models the float sum of its token weights in
 *                        token order; sum alpha is a float sum in topic order. */
int lda_topic_word_dist(lda_ctx* ctx, int topic, int normalize, float* out);
int lda_count_by_topic(lda_ctx* ctx, uint64_t* out);
int lda_doc_topic_dist(lda_ctx* ctx, int64_t doc, int normalize, float* out);

/* Streaming bandwidth of `device` (no counterpart in the reference; the
 * ceiling the sweep's bytes are set against): out[0] GB/s of a hand-written
 * streaming read of `bytes` (>= 1 MiB, multiple of 16) in 16-byte vectors,
 * out[1] GB/s of a copy (bytes read + bytes written), each over `reps`
 * launches after two warm-up launches, the best of four grid sizes.  Allocates 2 x bytes for the call. */
int lda_probe_bandwidth(int device, uint64_t bytes, int reps, double out[2]);

/* One sampling step per query token on the frozen current state: remove the
 * token's own z_old from (n_dk[doc], n_kw[:,word], n_k), weight, scan, and
 * search with the given uniform.  Counts are not modified. */
int lda_sample_frozen(lda_ctx* ctx, int64_t n, const int64_t* doc, const uint32_t* word, const uint16_t* z_old,
                      const float* u, uint16_t* z_new, float* total);

/* Fold-in of unseen documents against the current (frozen) n_kw / n_k.
 * mode 0: batched (one persistent launch over all documents, longest first,
 *         replayed from a cached hipGraph);
 * mode 1: plain (one launch and one synchronisation per document).
 * theta: D x K floats, ll: D doubles (FP64), z_out: N uint16 (may be NULL).
 * Models with word priors run tomoto's per-document sampler with eta_kv
 * (initial topics from the word's eta column, LDAModel.hpp:764-775); the
 * segmentation options apply to them as well, with at most the segment count
 * whose state fits the LDS at this K.  Term-weighted models need the token
 * weights: lda_infer_w. */
int lda_infer(lda_ctx* ctx, const uint32_t* words, const uint64_t* doc_off, int64_t D, int iterations,
              uint64_t infer_seed, int mode, float* theta, double* ll, uint16_t* z_out);

/* Batched fold-in that allocates no device memory (no counterpart in the
 * reference; for callers that cannot allocate: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same results as
 * lda_infer(mode 0) bit for bit; the token arena, jobs, parameters, bucket
 * counters and outputs of the call are staged in a 64 MiB device array that
 * the library holds once per GPU for the life of the process.  A call whose
 * arena (documents x 32 B + tokens x 6 B + documents x (4K + 8) B, each
 * region rounded up to 256 B, plus 512 B) exceeds 64 MiB fails with
 * LDA_ERR_INVALID_ARGUMENT.  Every call on one GPU shares that array, which
 * keeps the last call's contents: calls that overlap on one GPU (from
 * several threads or contexts) race and must be serialised by the caller.
 * TermWeight ONE models (with or without word priors), K <= 2048. */
int lda_infer_noalloc(lda_ctx* ctx, const uint32_t* words, const uint64_t* doc_off, int64_t D, int iterations,
                      uint64_t infer_seed, float* theta, double* ll, uint16_t* z_out);

/* Fold-in of a term-weighted (IDF / PMI) model: as lda_infer, with the
 * tokens' weights (tomoto computes them in initializeDocState,
 * LDAModel.hpp:794-809, from the model's vocabWeights: IDF w = vocabWeights[v],
 * PMI w = max(log(tf / vocabWeights[v] / len), 0)).  Each document is sampled
 * on a private copy of the float counts (adding w, removing max(x - w, 0)),
 * theta = (n_dk + alpha) / (sum of the weights + sum alpha) and the FP64 ll
 * uses the float counts.  Replaces ITopicModel::infer (TopicModel.hpp:842-875,
 * LDAModel.hpp:920-953) for these models.  weights must be NULL for
 * TermWeight ONE models (then this is lda_infer). */
int lda_infer_w(lda_ctx* ctx, const uint32_t* words, const uint64_t* doc_off, int64_t D, const float* weights,
                int iterations, uint64_t infer_seed, int mode, float* theta, double* ll, uint16_t* z_out);

/* Fold-in with together=True (LDAModel.hpp:852-892): the documents join a
 * copy of the model's n_kw / n_k (initial topics from Discrete(alpha), or with
 * word priors from the word's eta column, keyed by infer_seed) and are
 * sampled `iterations` times under the scheme and
 * workers of lda_train (partition chunks from the model's word frequencies),
 * with alpha fixed.  theta: D x K; *ll = getLLRest(copy) - getLLRest(model) +
 * getLLDocs(documents), FP64 (one value for all documents, as tomoto);
 * z_out: the documents' final topics (N, may be NULL). */
int lda_infer_together(lda_ctx* ctx, const uint32_t* words, const uint64_t* doc_off, int64_t D, int iterations,
                       uint64_t infer_seed, int workers, int scheme, float* theta, double* ll, uint16_t* z_out);

/* together=True for a term-weighted model: as lda_infer_together, with the
 * tokens' weights (see lda_infer_w).  The copy starts from the model's float
 * counts and the documents' tokens are added onto it in document and token
 * order (initializeDocState on tmpState, LDAModel.hpp:858-861); the sweeps
 * use the weighted schedules and merge; theta divides by the documents'
 * weight sums.  weights must be NULL for TermWeight ONE models. */
int lda_infer_together_w(lda_ctx* ctx, const uint32_t* words, const uint64_t* doc_off, int64_t D, const float* weights,
                         int iterations, uint64_t infer_seed, int workers, int scheme, float* theta, double* ll,
                         uint16_t* z_out);

/* Segmented fold-in (lda_infer and the server).  tomoto samples a document's
 * tokens one after the other (LDAModel.hpp:920-953), which bounds a long
 * document's latency by its length.  With seg_tokens > 0 a document of n >
 * seg_tokens tokens is split into S = min(seg_max, next power of two >=
 * ceil(n / seg_tokens)) segments sampled by S wavefronts; segment w starts
 * at the first word-run start at or after floor(n*w/S).  Each iteration every
 * segment samples against n_dk as it was when the iteration started plus its
 * own changes, and the changes are summed after the iteration.  Documents of
 * at most seg_tokens tokens, and every document when seg_tokens == 0 (the
 * default), are sampled exactly as tomoto does.  seg_max is 1, 2, 4 or 8 and
 * at most *seg_supported (what the kernel fits at this K). */
int lda_set_infer_options(lda_ctx* ctx, uint32_t seg_tokens, int seg_max);
int lda_get_infer_options(lda_ctx* ctx, uint32_t* seg_tokens, int* seg_max, int* seg_supported);

/* ---- Serving: continuous-batching fold-in over the context's frozen model ----
 * tomotopy has no server; these calls serve LDAModel.infer (py_LDA.cpp:345-411,
 * LDAModel::_infer<together=false>, LDAModel.hpp:920-953) to a stream of
 * requests.  A server keeps the context's n_kw / n_k on the device, admits
 * requests in FIFO order while their tokens fit a token arena and their
 * documents the active-document limit (both sized from hbm_budget_bytes), and
 * runs steps, each one persistent fold-in launch over the documents that are
 * waiting, on `lanes` (1..4) streams: a free lane starts a step as soon as
 * documents wait, while the other lanes' steps still run.  Per step a
 * document advances by min(step_iters, max(1, step_tokens / tokens per
 * wavefront)) iterations.  The server takes the context's segmentation
 * (lda_set_infer_options) at start.  Results equal lda_infer(mode 0) with the
 * same infer_seed, bit for bit (also with word priors).  TermWeight IDF / PMI
 * models and models with K above 2048 are served by the run-column fold-in:
 * each lane is a thread that folds the queued requests in as one batch, every
 * document running all its iterations in that step (step_iters / step_tokens
 * do not apply); their requests go through lda_serve_submit_w (IDF / PMI:
 * token weights as for lda_infer_w) and results equal lda_infer_w / lda_infer
 * (mode 0) with the same infer_seed, bit for bit.
 * The context must not be trained, re-uploaded or destroyed while a server on
 * it runs.  submit / wait may be called from any thread; lda_serve_stop must
 * not overlap a wait, and drops requests that were not waited for. */
typedef struct lda_server lda_server;

int lda_serve_start(lda_ctx* ctx, uint64_t hbm_budget_bytes, int iterations, int step_iters, int64_t step_tokens,
                    int lanes, uint64_t infer_seed, lda_server** out);
/* Copies the request (CSR as for lda_infer); returns its id at once. */
int lda_serve_submit(lda_server* server, const uint32_t* words, const uint64_t* doc_off, int64_t D,
                     uint64_t* request_id);
/* lda_serve_submit with token weights: requests to the server of a
 * TermWeight IDF / PMI model (weights[g] >= 0 and finite for valid words, as
 * LDAModel::infer takes them from the vocabulary weights,
 * LDAModel.hpp:794-809).  Replaces the same binding as lda_serve_submit. */
int lda_serve_submit_w(lda_server* s, const uint32_t* words, const uint64_t* doc_off, int64_t D, const float* weights,
                       uint64_t* request_id);

/* Waits for a request (timeout_ms < 0: no limit; on timeout LDA_ERR_STATE and
 * the request stays pending), copies theta (D x K) and ll (D), and the
 * latency from submit to results-on-host in ms; then forgets the request. */
int lda_serve_wait(lda_server* server, uint64_t request_id, int timeout_ms, float* theta, double* ll,
                   double* latency_ms);
/* stats: [0] steps, [1] device-step ms (unused, 0), [2] documents done,
 * [3] requests done, [4] token-iterations done, [5] active documents,
 * [6] queued requests, [7] token arena capacity, [8] total step time on the
 * device (ms, HIP events), [9] total host staging and launch time (ms),
 * [10] longest step on the device (ms), [11] mean documents per step. */
int lda_serve_stats(lda_server* server, double stats[12]);
int lda_serve_stop(lda_server* server);

/* Timing of the last lda_train / lda_sweep / lda_infer call, measured with
 * HIP events on the context's stream around the hot kernels only:
 * stats[0] total kernel ms, stats[1] launches, stats[2] tokens sampled,
 * stats[3] 1 when the call's device-wide sweeps ran on narrow columns. */
int lda_kernel_stats(lda_ctx* ctx, double stats[4]);

/* Narrow columns in the device-wide sweep of TermWeight ONE models (no
 * counterpart in the reference; on by default): for the duration of an
 * lda_train / lda_sweep call, the n_kw columns of words with fewer than 2^15
 * tokens are kept as u16 (after the last word with more, in word-id order:
 * prepared vocabularies are sorted by frequency, so that is every word past
 * the most frequent few), halving the bytes the sweep reads per token of
 * those words; they are written back to the int32 counts at the end of the
 * call.  Counts outside [0, 2^15) in those columns (set from outside and not
 * matching the corpus) make the call run on int32 columns.  enable = 0 turns
 * the narrow columns off. */
int lda_set_narrow_columns(lda_ctx* ctx, int enable);

/* Tokens per group of shared n_kw atomics in the device-wide sweeps with K <=
 * 2048 and K above it (no counterpart in the reference): each wavefront sends
 * the n_kw changes of a group of its tokens at the group's end, as one atomic
 * per token from the lane that holds it, instead of one single-lane atomic per
 * token, which held the wavefront's next loads behind the atomic's round trip.
 * A wavefront never rereads a column with changes pending, so one wavefront
 * alone samples the same chain for every group; other workers see the changes
 * up to one group later.  tokens = 0 (the default) picks the largest power of
 * two up to 32 that keeps the changes held back by all wavefronts at most
 * 1/128 of the corpus's tokens; otherwise a power of two up to 64. */
int lda_set_atomic_group(lda_ctx* ctx, int tokens);

#ifdef __cplusplus
}
#endif

#endif /* TOMOTOPY_AMD_H */
