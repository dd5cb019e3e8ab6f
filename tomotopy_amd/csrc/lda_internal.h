// Internal interface between the host core (lda_capi.cpp) and the gfx950
// kernels (lda_device.hip).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lda {

// Philox stream tags (4th counter word); the CPU oracle uses the same values.
constexpr uint32_t TAG_SWEEP = 0x53574550u;
constexpr uint32_t TAG_INIT = 0x494E4954u;
constexpr uint32_t TAG_SHUFFLE = 0x53484646u;
constexpr uint32_t TAG_INFER = 0x494E4652u;
constexpr uint32_t TAG_INFINIT = 0x49494E49u;

// Device K-vector layout ("lane layout").  A K-vector is padded to
// Kp = 64*kpl entries; lane l of a wavefront owns topics [l*kpl, (l+1)*kpl).
// For kpl >= 4 the entries are stored so that the g-th 16-byte load of every
// lane is one contiguous 1 KiB segment: topic k = l*kpl + j lives at
// (j/4)*256 + l*4 + j%4.  For kpl < 4 the layout is the natural order.
inline int kvec_pos(int kpl, int k) {
    if (kpl < 4) return k;
    int l = k / kpl, j = k % kpl;
    return (j >> 2) * 256 + l * 4 + (j & 3);
}

struct SweepParams {
    const uint32_t* words;
    uint16_t* z;
    const uint64_t* off;      // D+1 document offsets
    int32_t* ndk;             // D x Kp
    int32_t* nkw;             // V x Kp
    const int32_t* nk_in;     // device-wide: n_k snapshot; det: private n_k (W x Kp)
    int32_t* nk_out;          // device-wide: accumulates deltas; det: private n_k (same as nk_in)
    const float* alpha;       // Kp, zero padded
    float eta, veta;
    int K, V, Kp;
    uint32_t seed_lo, seed_hi, sweep;
    int64_t D;
    const uint32_t* doc_order;  // device-wide: documents, longest first
    uint32_t* counter;          // device-wide: next document
    const uint64_t* coff;       // det: (W+1) x D chunk offsets
    int W, round;
    const float* wt;            // term-weighted models: token weights (counts are float bits)
    // word priors (NULL prior_row: symmetric eta): prior_row[v] is the row of
    // word v's eta column in prior_eta (rows of Kp, lane layout) or -1;
    // etasum = sum_v eta_kv per topic (Kp, lane layout, 1 in padded topics)
    const int32_t* prior_row;
    const float* prior_eta;
    const float* etasum;
    // partition rounds replayed from a graph: the sweep number is read here
    // (NULL: use `sweep`)
    const uint32_t* sweep_dev;
    // narrow columns of the device-wide sweep (NULL: every column in nkw):
    // words v >= nkw_hot keep their counts as u16 in nkw16, row v - nkw_hot
    // (same lane layout); their int32 columns in nkw are stale meanwhile
    uint16_t* nkw16;
    int nkw_hot;
    // K above 2048 (kpl > kSmallMaxKpl), device-wide sweep: per worker two
    // rows of Kp counts (its private n_k and its copy of the current column)
    void* bk_scratch;
    // integer counts with a document over 2^20 tokens or an alpha >= 2^20:
    // the K <= 2048 kernels keep float(n_dk) + alpha and recover n_dk from it,
    // exact only below 2^21, so these run on the lda_bigk.hip kernels (n_dk
    // kept as integers) at any K
    int exact;
    // device-wide sweeps (K <= 2048): tokens per group of shared n_kw atomics
    // (a power of two up to 64; a wavefront holds back at most that many
    // tokens' n_kw changes)
    int col_group;
    // device-wide sweeps with K above 2048 in workgroups of several
    // wavefronts: each wavefront's n_dk row (bit 0) and column copy (bit 1)
    // in LDS as well
    int bk_rows;
};

// Topics per lane: the register / LDS kernels take kpl <= kSmallMaxKpl
// (K <= 2048); lda_bigk.hip takes kSmallMaxKpl < kpl <= kBigMaxKpl
// (K <= 32768, tomotopy's limit being 32767).
constexpr int kSmallMaxKpl = 32;
constexpr int kBigMaxKpl = 512;

// Words with fewer than kNarrowCf tokens have every n_kw cell below 2^15
// while the counts are consistent, and a sweep can raise a cell by at most
// the word's token count, so u16 cells cannot overflow.
constexpr uint64_t kNarrowCf = 1u << 15;

// One fold-in job: a document's tokens [b, e) in the token arena, the
// iterations (it_begin, it_end] to run in this launch out of `iters`, and the
// output slot of theta / ll.  it_begin == 0 draws the initial topics;
// it_end == iters writes theta and the log-likelihood.
struct InferJob {
    uint64_t b, e;
    int32_t it_begin, it_end, iters, slot;
};

// Segmented fold-in.  A document of n tokens is split into S segments, one
// wavefront each, S = 1 when seg_tokens == 0 or n <= seg_tokens, else the
// smallest power of two >= ceil(n / seg_tokens), at most seg_max.  Segment
// w starts at the first word-run start at or after floor(n*w/S), so every
// word's tokens lie in one segment.  Within an iteration each segment samples
// its tokens in order against n_dk as it was when the iteration started plus
// its own changes; the segments' changes are summed after every iteration.
// S == 1 is the sequential sampler.  Jobs are grouped by S into buckets
// S = 1, 2, 4, 8 (bucket log2 S), one launch each.
constexpr int kFoldBuckets = 4;
constexpr int kMaxFoldSegments = 8;

inline int fold_segments(uint64_t n, uint32_t seg_tokens, int seg_max) {
    if (seg_tokens == 0 || n <= seg_tokens) return 1;
    const uint64_t need = (n + seg_tokens - 1) / seg_tokens;
    int S = 1;
    while ((uint64_t)S < need && S < seg_max) S <<= 1;
    return S;
}

inline int fold_bucket(int S) { return S >= 8 ? 3 : S >= 4 ? 2 : S >= 2 ? 1 : 0; }

struct InferParams {
    const uint32_t* words;   // token arena
    uint16_t* z;             // token arena
    const InferJob* jobs;    // grouped by bucket
    int64_t njobs;
    uint32_t* counters;      // next job per bucket; NULL: the launch runs job job_base only
    int64_t job_base;
    int64_t bucket_lo[kFoldBuckets], bucket_hi[kFoldBuckets];  // job index range of each bucket
    const int32_t* nkw;      // frozen V x Kp
    const int32_t* nk;       // frozen Kp
    const float* alpha;      // Kp
    const float* alpha_cdf;  // K, natural order, sequential float sum
    float eta, veta;
    int K, V, Kp;
    uint32_t seed_lo, seed_hi;
    float* theta;            // slots x K natural order
    double* ll;              // slots
    double asum_d;           // sum of alpha in double
    float asum_f;            // sequential float sum of alpha
    // word priors (the ASYM kernels): prior_row[v] = row of word v's eta
    // column in prior_eta (Kp, lane layout) or -1; per-topic sums of eta_kv
    const int32_t* prior_row;
    const float* prior_eta;
    const float* etasum;
    const double* etasum_d;
};

// Fold-in of term-weighted models (wt set: the count buffers hold float32
// bits) and of models with word priors (prior_row set).  Jobs as InferJob
// (it_begin 0, it_end == iters); the runs of equal word ids of job i are
// [job_run[i], job_run[i+1]) in token order, each with a private column of
// Kp cells in `cells`.
struct InferGenParams {
    const uint32_t* words;
    uint16_t* z;
    const float* wt;
    const InferJob* jobs;
    int64_t njobs;
    uint32_t* counter;           // next job; NULL: the launch runs job job_base only
    int64_t job_base;
    const int64_t* job_run;      // njobs + 1
    // segmented documents (S = workgroup size / 64 > 1): per job, rows of
    // kMaxFoldSegments + 1 entries: segment w's first token (seg_b[w], w <= S)
    // and the index of its first run (seg_r[w])
    const uint64_t* seg_b;
    const int64_t* seg_r;
    const uint32_t* run_word;
    int32_t* cells;              // runs x Kp
    int32_t* ndk_out;            // slots x Kp, lane layout
    int32_t* nk_out;             // slots x Kp, lane layout
    float* docw;                 // slots: the document's weight sum (token count for integer counts)
    const int32_t* nkw;          // frozen V x Kp
    const int32_t* nk;           // frozen Kp
    const float* alpha;          // Kp, lane layout
    const float* alpha_cdf;      // K, natural order
    float eta, veta;
    double eta_d, veta_d;        // double(eta), V * double(eta)
    int K, V, Kp;
    const int32_t* prior_row;
    const float* prior_eta;
    const float* etasum;
    const double* etasum_d;
    uint32_t seed_lo, seed_hi;
    float* theta;                // slots x K, natural order
    double* ll;                  // slots
    double asum_d;
    float asum_f;
    // the finalize's run pass: chunks of at most kGenChunkRuns runs of one
    // document; chunk_base (njobs + 1) is each job's first chunk, chunk_rb
    // (chunks + 1) each chunk's first run, chunk_ll (chunks) its n_kw terms
    const int64_t* chunk_base;
    const int64_t* chunk_rb;
    double* chunk_ll;
    // integer counts with a document over 2^20 tokens or an alpha >= 2^20
    // (see SweepParams::exact): the lda_bigk.hip fold-in at any K
    int exact;
};
constexpr int64_t kGenChunkRuns = 128;
// finalize = false: the sampler (grid persistent workgroups of one
// wavefront); true: theta and the log-likelihood (the run pass over the
// jobs' chunks, then the document pass, each with grid wavefronts).
// segs: wavefronts per document (1, 2, 4, 8) of the sampler launch.
hipError_t launch_infer_gen(int kpl, int segs, const InferGenParams& p, int grid, bool finalize, hipStream_t s);
int infer_gen_blocks_per_cu(int kpl, bool wt, bool asym, int segs);

// Term-weighted (IDF / PMI) training: the count buffers hold float32 values.
hipError_t launch_sweep_device_w(int kpl, const SweepParams& p, int grid, int waves, hipStream_t s);
hipError_t launch_sweep_det_w(int kpl, const SweepParams& p, hipStream_t s);
int sweep_device_w_blocks_per_cu(int kpl, bool asym, int waves);
// copy_merge on float counts: W private copies as in launch_sweep_copy; the
// merge is n_kw = local_0 + sum_{w>=1} (local_w - old) in worker order
hipError_t launch_sweep_copy_w(int kpl, const SweepParams& p, hipStream_t s);
hipError_t launch_merge_copies_f(float* nkw, const float* loc, int64_t n, int W, hipStream_t s);
// n_kw = max(n_kw, 0), n_k = rowwise sum of n_kw, in float (LDAModel.hpp:565-569)
// Words per block of the weighted merge's two-level row sum (n_k from n_kw).
constexpr int kRowsumBlock = 256;
// part: ceil(V / kRowsumBlock) * Kp floats of scratch.
hipError_t launch_clamp_rowsum_f(float* nkw, int V, int Kp, float* part, float* nk, hipStream_t s);
hipError_t launch_loglik_f(const float* ndk, const float* docw, int64_t D, const float* nkw, const float* nk,
                           const float* alpha, int V, int Kp, double asum, double eta, double veta, double* partial,
                           int nblk, const int32_t* prior_row, const float* prior_eta, const double* etasum_d,
                           hipStream_t s);
hipError_t launch_alpha_sums_f(const float* ndk, const float* docw, int64_t D, const float* alpha, int Kp,
                               const float* asum, double* part, int nblk, hipStream_t s);

// kernels (lda_device.hip)
// z ~ Discrete(alpha), or with word priors (prior_row set) z ~ the word's eta column
// recip_rn against IEEE division over all 2^23 mantissas of the binade
// [2^exp_shift, 2^(exp_shift+1)); *bad (zeroed by the caller) counts mismatches.
hipError_t launch_selftest_reciprocal(int exp_shift, unsigned long long* bad, hipStream_t s);

hipError_t launch_init_topics(const uint32_t* words, uint16_t* z, uint64_t N, const float* alpha_cdf, int K, int V,
                              uint32_t seed_lo, uint32_t seed_hi, const int32_t* prior_row, const float* prior_eta,
                              int kpl, int Kp, float eta, hipStream_t s);
hipError_t launch_build_counts(int kpl, const uint32_t* words, const uint16_t* z, const uint64_t* off, int64_t D,
                               int32_t* ndk, int32_t* nkw, int32_t* nk, int Kp, int V, hipStream_t s);
// Device-wide sweep: grid workgroups (workers) of `waves` wavefronts (1..8)
// sharing one private n_k.
constexpr int kMaxSweepWaves = 8;
hipError_t launch_sweep_device(int kpl, const SweepParams& p, int grid, int waves, hipStream_t s);
// n_kw rows [hot, V) between int32 (nkw) and u16 (nkw16, row v - hot); pack
// sets *bad when a cell is outside [0, kNarrowCf)
hipError_t launch_pack_narrow(const int32_t* nkw, uint16_t* nkw16, int hot, int V, int Kp, uint32_t* bad, hipStream_t s);
// Read-outs: one topic's (n_kw + eta) / sum over the words (fl: float
// counts); tokens per topic over the valid tokens (cnt zeroed by the caller).
hipError_t launch_topic_word_dist(const void* nkw, bool fl, int V, int Kp, int pos, float eta, float sum, int normalize,
                                  float* out, hipStream_t s);
hipError_t launch_count_topics(const uint32_t* words, const uint16_t* z, uint64_t N, int V, int K,
                               unsigned long long* cnt, hipStream_t s);
// Bandwidth probes: a streaming read of `bytes` (multiple of 16) and a copy.
hipError_t launch_probe_read(const void* src, int64_t bytes, int* sink, int grid, hipStream_t s);
hipError_t launch_probe_copy(const void* src, void* dst, int64_t bytes, int grid, hipStream_t s);
hipError_t launch_unpack_narrow(int32_t* nkw, const uint16_t* nkw16, int hot, int V, int Kp, hipStream_t s);
hipError_t launch_sweep_det(int kpl, const SweepParams& p, hipStream_t s);
// copy_merge: W private copies of n_kw (W x V x Kp) and of n_k (W x Kp)
hipError_t launch_sweep_copy(int kpl, const SweepParams& p, hipStream_t s);
hipError_t launch_merge_copies(int32_t* nkw, const int32_t* loc, int64_t n, int W, hipStream_t s);
hipError_t launch_replicate(int32_t* dst, const int32_t* src, int64_t n, int W, hipStream_t s);
int sweep_device_blocks_per_cu(int kpl, bool asym, int waves);
// Resident one-wavefront workgroups per CU of the partition round kernel
// (integer or float counts).
int sweep_det_blocks_per_cu(int kpl, bool asym, bool weighted);
hipError_t launch_copy_i32(int32_t* dst, const int32_t* src, int64_t n, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* dst, uint32_t v, int64_t n, hipStream_t s);
hipError_t launch_add_i32(int32_t* dst, const int32_t* src, int64_t n, hipStream_t s);
hipError_t launch_rowsum_nkw(const int32_t* nkw, int V, int Kp, int32_t* nk, hipStream_t s);

// FP64 log-likelihood partial sums: out[0..nblk) docs part, out[nblk..2nblk) rest part.
// prior_row / prior_eta / etasum_d: word priors (NULL: symmetric eta).
hipError_t launch_loglik(const int32_t* ndk, const int32_t* doclen, int64_t D, const int32_t* nkw, const int32_t* nk,
                         const float* alpha, int V, int Kp, double asum, double eta, double veta, double* partial,
                         int nblk, const int32_t* prior_row, const float* prior_eta, const double* etasum_d,
                         hipStream_t s);

// alpha fixed point: per-topic sums of digammaApprox(n_dk + alpha_k) and the
// denominator sum of digammaApprox(L_d + asum); part[b*(Kp+1) + pos].
hipError_t launch_alpha_sums(const int32_t* ndk, const int32_t* doclen, int64_t D, const float* alpha, int Kp,
                             const float* asum, double* part, int nblk, hipStream_t s);
// The alpha fixed point on the device (optimizeParameters, LDAModel.hpp:221-265).
// asum: device float, the sum of alpha in topic order (launch_alpha_asum);
// dglut: tomoto's digammaT table (kDigammaLutSize floats); bad: set when an
// alpha leaves [1e-5, amax).  Direct: from alpha_sums partials.  Histogram
// (integer counts): hist = Kp x (Lmax + 1) and hist_len = Lmax + 1 zeroed ints.
constexpr int kDigammaLutSize = 1024 + 6400 + 4000;
hipError_t launch_alpha_asum(const float* alpha, int K, int kpl, float* asum, hipStream_t s);
hipError_t launch_alpha_update_direct(float* alpha, const double* part, int nblk, int Kp, int64_t D, const float* asum,
                                      const float* dglut, float amax, unsigned* bad, hipStream_t s);
hipError_t launch_alpha_hist(const int32_t* ndk, const int32_t* doclen, int64_t D, int Kp, int Lmax, int32_t* hist,
                             int32_t* hist_len, hipStream_t s);
hipError_t launch_alpha_den_hist(const int32_t* hist_len, int Lmax, int64_t D, const float* asum, const float* dglut,
                                 float* den, hipStream_t s);
hipError_t launch_alpha_update_hist(float* alpha, const int32_t* hist, int Lmax, int Kp, int64_t D, const float* den,
                                    const float* dglut, float amax, unsigned* bad, hipStream_t s);
hipError_t launch_doclen(const uint32_t* words, const uint64_t* off, int64_t D, int V, int32_t* doclen, hipStream_t s);

// invariants
hipError_t launch_invariants(int kpl, const uint32_t* words, const uint16_t* z, const uint64_t* off, int64_t D,
                             const int32_t* ndk, const int32_t* nkw, const int32_t* nk, const int32_t* doclen,
                             int K, int V, int Kp, int32_t* scratch_ndk, int32_t* scratch_nkw,
                             long long* colsum, unsigned long long* flags, hipStream_t s);

// frozen-state single steps
// bk_scratch: 3 * Kp counts per workgroup of bk_grid (kpl > kSmallMaxKpl only)
hipError_t launch_sample_frozen(int kpl, const int32_t* ndk, const int32_t* nkw, const int32_t* nk,
                                const float* alpha, float eta, float veta, int K, int V, int Kp, int64_t n,
                                const int64_t* doc, const uint32_t* word, const uint16_t* z_old, const float* u,
                                uint16_t* z_new, float* total, int32_t* bk_scratch, int bk_grid, hipStream_t s);

// fold-in; the kernel reads its parameters from device memory so that one
// captured graph serves every batch
// segs: wavefronts per document (1, 2, 4 or 8), the bucket this launch serves
// asym: the word-prior variant (InferParams.prior_row set)
hipError_t launch_infer(int kpl, int segs, bool asym, const InferParams* p_dev, int grid, hipStream_t s);
// theta and FP64 log-likelihood of the jobs that finish (it_end == iters)
hipError_t launch_infer_finalize(int kpl, bool asym, const InferParams* p_dev, int grid, hipStream_t s);
// resident workgroups per CU of the fold-in kernel (0: this segment count does not fit)
int infer_blocks_per_cu(int kpl, int segs, bool asym);

// K above 2048 (lda_bigk.hip): one wavefront per worker or document, state in
// device memory.  Device-wide: grid workers, p.bk_scratch holds 2 * Kp counts
// per worker.  Frozen steps: scratch holds 3 * Kp counts per workgroup of grid.
// Fold-in: one wavefront per document, no segmentation.
hipError_t launch_bk_sweep_det(int kpl, const SweepParams& p, bool weighted, hipStream_t s);
hipError_t launch_bk_sweep_copy(int kpl, const SweepParams& p, bool weighted, hipStream_t s);
// waves == 1: one-wavefront workers, n_k and alpha read from device memory;
// waves > 1: workgroups of `waves` wavefronts sharing alpha and n_k in LDS
// (bk_scratch then holds one column row per wavefront)
hipError_t launch_bk_sweep_device(int kpl, const SweepParams& p, bool weighted, int grid, int waves, hipStream_t s);
size_t bk_sweep_device_lds_bytes(int kpl, int Kp, int waves, int rows);
int bk_sweep_device_lds_blocks_per_cu(bool weighted, bool asym, int kpl, int Kp, int waves, int rows);
int bk_sweep_blocks_per_cu(bool weighted, bool asym, bool device, int kpl);
hipError_t launch_bk_sample_frozen(int kpl, const int32_t* ndk, const int32_t* nkw, const int32_t* nk,
                                   const float* alpha, float eta, float veta, int K, int Kp, int64_t n, int grid,
                                   const int64_t* doc, const uint32_t* word, const uint16_t* z_old, const float* u,
                                   uint16_t* z_new, float* total, int32_t* scratch, hipStream_t s);
hipError_t launch_bk_infer_gen(int kpl, const InferGenParams& p, int grid, bool finalize, hipStream_t s);
int bk_infer_gen_blocks_per_cu(bool wt, bool asym, int kpl);

}  // namespace lda
