// ============================================================================
// lda_oracle.cpp -- CPU ORACLE for the collapsed-Gibbs LDA hot path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load this library, and only as the checker
// (or as the timed CPU baseline).  The product (tomotopy_amd) never links,
// loads or calls it; the product path fails loudly without its HIP library.
//
// This is a plain C++17 restatement of tomoto's plain-LDA Gibbs sampler
// (bab2min/tomotopy, src/TopicModel/LDAModel.hpp and friends), written from
// the reference's behaviour, not copied.  Every function cites the reference
// file:line whose semantics it restates.  Two deliberate substitutions:
//   * RNG: tomoto draws from EigenRand's P8_mt19937_64_32 (TopicModel.hpp:16);
//     EigenRand 0.5.1 is absent here, so both this oracle and the device use
//     Philox4x32-10 (Salmon et al., SC'11) with a counter layout defined in
//     lda_rng_counter() below.  The per-token uniform is a pure function of
//     (seed, sweep, token index), so results do not depend on schedule.
//   * Scan order: tomoto's prefixSum uses an SSE 4-block association
//     (sample.hpp:61-86).  The oracle implements that order AND the device
//     order (lane-local sequential scan + 64-lane DPP row scan); the device
//     is checked bit-exactly against the device order, and the tie rate of
//     device order vs tomoto order is measured.
//
// Parity pins (see tests/): Philox known-answer vectors (Random123 KAT),
// weights vs numpy float32 closed form, scans vs numpy emulations, search vs
// brute force, count invariants, FP64 LL vs scipy gammaln, tomoto LUT lgamma
// vs FP64 within its interpolation bound, alpha fixed point on Dir(alpha)
// data, infer delta-LL vs brute-force full-copy LL.
//
// Build: oracle/Makefile -> oracle/_build/liblda_oracle.so (g++ -O2
// -ffp-contract=off; SSE float math, no x87, no FMA contraction).
// ============================================================================
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------- Philox ----
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11).  Constants are the published ones.
constexpr uint32_t PHILOX_M0 = 0xD2511F53u, PHILOX_M1 = 0xCD9E8D57u;
constexpr uint32_t PHILOX_W0 = 0x9E3779B9u, PHILOX_W1 = 0xBB67AE85u;

inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += PHILOX_W0; k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Stream tags (4th counter word).  Shared with the device code.
enum : uint32_t {
    TAG_SWEEP = 0x53574550u,   // per-token uniform of a training sweep
    TAG_INIT = 0x494E4954u,    // initial topic draw (LDAModel.hpp:754-762)
    TAG_SHUFFLE = 0x53484646u, // forShuffled seed (Utils.hpp:157-178)
    TAG_INFER = 0x494E4652u,   // fold-in sweep uniform (LDAModel.hpp:930-940)
    TAG_INFINIT = 0x49494E49u, // fold-in initial draw
};

inline uint32_t philox_word(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t tag) {
    uint32_t ctr[4] = {c0, c1, c2, tag};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t out[4];
    philox4x32_10(ctr, key, out);
    return out[0];
}

// u in [0,1) with 24 random bits: exact in float.
inline float word_to_uniform(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

inline float token_uniform(uint64_t seed, uint64_t g, uint32_t sweep, uint32_t tag) {
    return word_to_uniform(philox_word(seed, (uint32_t)g, (uint32_t)(g >> 32), sweep, tag));
}

// ------------------------------------------------------- sampling pieces ----
// a1: zLikelihood[k] = (n_dk+alpha_k) * (n_kw+eta) / (n_k + V*eta)
// (LDAModel.hpp:273-285, EtaHelper 81-97: V*eta is float(eta)*float(realV)).
inline float topic_weight(int32_t ndk, float alpha, int32_t nkw, float eta, int32_t nk, float veta) {
    float a = (float)ndk + alpha;
    float b = (float)nkw + eta;
    float num = a * b;
    float den = (float)nk + veta;
    return num / den;
}

// a2, tomoto order: sample.hpp:61-86.  Blocks of 4 are scanned as
// [a, a+b, (b+c)+a, (c+d)+(a+b)] and offset by the previous block's last
// element; the tail (or everything when n < 4, starting at index 1) is a
// sequential running sum.
void prefix_sum_tomoto(float* x, int n) {
    int n4 = n & ~3;
    float offset = 0.0f;
    for (int i = 0; i < n4; i += 4) {
        float a = x[i], b = x[i + 1], c = x[i + 2], d = x[i + 3];
        float ab = a + b, bc = b + c, cd = c + d;
        float o0 = a, o1 = ab, o2 = bc + a, o3 = cd + ab;
        if (i) { o0 = o0 + offset; o1 = o1 + offset; o2 = o2 + offset; o3 = o3 + offset; }
        x[i] = o0; x[i + 1] = o1; x[i + 2] = o2; x[i + 3] = o3;
        offset = o3;
    }
    int start = n4 ? n4 : 1;
    for (int i = start; i < n; ++i) x[i] = x[i] + x[i - 1];
}

// a2, device order.  Topics are laid out lane-major: lane l of the 64-lane
// wavefront owns topics [l*kpl, (l+1)*kpl).  Each lane scans its own topics
// sequentially; lane totals are scanned with the DPP pattern row_shr:1,2,4,8
// (zero fill outside the 16-lane row), then row_bcast:15 into rows 1 and 3,
// then row_bcast:31 into rows 2 and 3; the exclusive lane prefix is the
// inclusive value of lane l-1 (wave_shr:1, zero into lane 0) and is added to
// every lane-local partial.  x has 64*kpl entries (padding entries are 0).
void prefix_sum_device(float* x, int kpl) {
    thread_local std::vector<float> part_buf;
    part_buf.resize((size_t)64 * kpl);
    float* part = part_buf.data();
    float tot[64];
    for (int l = 0; l < 64; ++l) {
        float s = x[l * kpl];
        part[l * kpl] = s;
        for (int j = 1; j < kpl; ++j) { s = s + x[l * kpl + j]; part[l * kpl + j] = s; }
        tot[l] = s;
    }
    float v[64], t[64];
    std::memcpy(v, tot, sizeof v);
    for (int d = 1; d <= 8; d <<= 1) {
        std::memcpy(t, v, sizeof v);
        for (int l = 0; l < 64; ++l) if ((l & 15) >= d) v[l] = t[l] + t[l - d];
    }
    std::memcpy(t, v, sizeof v);
    for (int l = 16; l < 32; ++l) v[l] = t[l] + t[15];
    for (int l = 48; l < 64; ++l) v[l] = t[l] + t[47];
    std::memcpy(t, v, sizeof v);
    for (int l = 32; l < 64; ++l) v[l] = t[l] + t[31];
    for (int l = 0; l < 64; ++l) {
        float e = l ? v[l - 1] : 0.0f;
        for (int j = 0; j < kpl; ++j) x[l * kpl + j] = e + part[l * kpl + j];
    }
}

// a3: sampleFromDiscreteAcc (sample.hpp:133-189): r = u * cdf[K-1]; return
// the first z with r < cdf[z], K-1 if there is none.
inline int search_cdf(const float* cdf, int K, float u) {
    float r = u * cdf[K - 1];
    for (int z = 0; z < K; ++z) if (r < cdf[z]) return z;
    return K - 1;
}

// sampleFromDiscrete (sample.hpp:121-131), the initial draw with word priors
// (LDAModel.hpp:769-773): r = u * (sequential float sum of w), then subtract
// w[z] while r > w[z] and z < K-1.
int sample_from_discrete(const float* w, int K, float u) {
    float tot = 0.0f;
    for (int k = 0; k < K; ++k) tot += w[k];
    float r = u * tot;
    int z = 0;
    while (r > w[z] && z < K - 1) { r -= w[z]; ++z; }
    return z;
}

// kpl used by the device for K topics: the smallest power of two with 64*kpl >= K.
inline int device_kpl(int K) {
    int kpl = 1;
    while (64 * kpl < K) kpl <<= 1;
    return kpl;
}

// Weight vector + scan + search, the full per-token step of a1-a3.  When
// scan_order == 1 the device order is used (weights padded to 64*kpl with 0).
// eta_col / vetas: the word-prior variant (EtaHelper<_, true>,
// LDAModel.hpp:99-117): eta_kv of this word and sum_v eta_kv per topic.
int sample_one(int K, int scan_order, int kpl, const int32_t* ndk, const float* alpha,
               const int32_t* nkw_col, float eta, const int32_t* nk, float veta, float u, float* buf,
               const float* eta_col = nullptr, const float* vetas = nullptr) {
    int n = scan_order ? 64 * kpl : K;
    for (int k = 0; k < K; ++k)
        buf[k] = topic_weight(ndk[k], alpha[k], nkw_col[k], eta_col ? eta_col[k] : eta, nk[k], vetas ? vetas[k] : veta);
    for (int k = K; k < n; ++k) buf[k] = 0.0f;
    if (scan_order) prefix_sum_device(buf, kpl);
    else prefix_sum_tomoto(buf, K);
    return search_cdf(buf, K, u);
}

// forShuffled (Utils.hpp:157-178): visit 0..N-1 in the order ((i+seed)*P) % N,
// P a prime from a fixed table chosen by seed (stepping when it divides N).
void shuffled_order(size_t N, uint32_t seed, std::vector<uint32_t>& out) {
    static const size_t primes[16] = {65537, 65539, 65543, 65551, 65557, 65563, 65579, 65581,
                                      65587, 65599, 65609, 65617, 65629, 65633, 65647, 65651};
    out.resize(N);
    if (!N) return;
    size_t P = primes[seed & 0xF];
    if (N % P == 0) {
        P = primes[(seed + 1) & 0xF];
        if (N % P == 0) P = primes[(seed + 2) & 0xF];
        if (N % P == 0) P = primes[(seed + 3) & 0xF];
    }
    P %= N;
    for (size_t i = 0; i < N; ++i) out[i] = (uint32_t)(((i + (size_t)seed) * P) % N);
}

// -------------------------------------------------------- lgamma LUT ------
// tomoto's float LUT lgamma / digamma (LUT.hpp:10-79, math.h:117-154).
// The lgamma grid is C lgamma(double) at float grid points rounded to float,
// and its small/large tails mix float and double arithmetic, as tomoto's
// unqualified lgamma/log calls on floats resolve to the C double functions;
// the digamma grid is a double-precision digamma rounded to float (tomoto uses
// its own digamma series).  The interpolation structure (steps 1/1024, 1/64,
// 1/8 or 1/4) and the small/large-argument tails are the reference's.
double digamma_d(double x) {
    double r = 0.0;
    while (x < 6.0) { r -= 1.0 / x; x += 1.0; }
    double f = 1.0 / (x * x);
    double t = f * (-1.0 / 12 + f * (1.0 / 120 + f * (-1.0 / 252 + f * (1.0 / 240 + f * (-1.0 / 132)))));
    return r + std::log(x) - 0.5 / x + t;
}

struct Lut3 {
    int N, M, L; float P, Q, R;
    int kind;  // 0 lgamma, 1 digamma
    std::vector<float> pts;
    Lut3(int kind_, int U) : N(1024), M(6400), L(kind_ ? 4000 : 8000), kind(kind_) {
        P = 1.0f / 1024; Q = 1.0f / 64; R = 1.0f / (float)U;
        pts.resize(N + M + L);
        for (int i = 0; i < N; ++i) pts[i] = f(i ? i * P : 0.0001f);
        for (int i = 0; i < M; ++i) pts[N + i] = f(i * Q + N * P);
        for (int i = 0; i < L; ++i) pts[N + M + i] = f(i * R + N * P + M * Q);
    }
    float f(float x) const { return kind ? (float)digamma_d((double)x) : (float)::lgamma((double)x); }
    float small(float x) const {
        const float pi = 3.14159265f;
        if (kind == 0) {
            if (x == 0) return INFINITY;
            return (float)((double)(x + 0.5f) * ::log((double)(x + 1)) - (double)(x + 1) + ::log((double)(2 * pi)) / 2 +
                           (double)(1 / 12.f / (x + 1)) - ::log((double)x));
        }
        if (x == 0) return -INFINITY;
        return logf(x + 2) - 0.5f / (x + 2) - 1 / 12.f / powf(x + 2, 2) - 1 / (x + 1) - 1 / x;
    }
    float large(float x) const {
        const float pi = 3.14159265f;
        if (kind == 0)
            return (float)((double)(x - 0.5f) * ::log((double)x) - (double)x + ::log((double)(2 * pi)) / 2 +
                           (double)(1 / 12.f / x));
        return logf(x) - 0.5f / x - 1 / 12.f / powf(x, 2);
    }
    float get(float x) const {
        if (!std::isfinite(x)) {
            if (std::isnan(x)) return NAN;
            if (kind == 0) return x > 0 ? INFINITY : -INFINITY;
            return x < 0 ? NAN : INFINITY;
        }
        if (x < 0) return NAN;
        if (x < 0.001f) return small(x);
        if (x >= N * P + M * Q + (L - 1) * R) return large(x);
        size_t idx; float a; float nx = x;
        if (x < N * P) { idx = (size_t)(nx / P); a = (nx - idx * P) / P; }
        else {
            nx -= N * P;
            if (nx < M * Q) { idx = (size_t)(nx / Q); a = (nx - idx * Q) / Q; idx += N; }
            else { nx -= M * Q; idx = (size_t)(nx / R); a = (nx - idx * R) / R; idx += N + M; }
        }
        return pts[idx] + a * (pts[idx + 1] - pts[idx]);
    }
};

const Lut3& lut_lgamma() { static Lut3 t(0, 8); return t; }
const Lut3& lut_digamma() { static Lut3 t(1, 4); return t; }

// digammaApprox (math.h:170-175), evaluated in float as tomoto does over
// float arrays.
inline float digamma_approx_f(float z) {
    return logf(z + 4) - 0.5f / (z + 4) - (1.0f / 12.0f) / ((z + 4) * (z + 4)) - 1.0f / z - 1.0f / (z + 1) -
           1.0f / (z + 2) - 1.0f / (z + 3);
}

// ------------------------------------------------------------- model ------
struct Model {
    int K, V, kpl, scan_order;
    float eta, veta;
    uint64_t seed;
    std::vector<float> alpha;
    // corpus (words sorted ascending per document; ids >= V are skipped)
    std::vector<uint32_t> words;
    std::vector<uint64_t> off;
    int64_t D = 0;
    std::vector<uint16_t> z;
    std::vector<int32_t> ndk;  // D x K
    std::vector<int32_t> nkw;  // V x K (word-major, tomoto's column-major K x V)
    std::vector<int32_t> nk;   // K
    // word priors (set_word_prior): eta_kv (V x K) and its per-topic sums, in
    // float (sequential over words, the sampler's) and double (the FP64 LL's);
    // empty without priors
    std::vector<float> eta_kv, etasum;
    std::vector<double> etasum_d;
    const float* ecol(uint32_t v) const { return eta_kv.empty() ? nullptr : &eta_kv[(size_t)v * K]; }
    const float* esum() const { return etasum.empty() ? nullptr : etasum.data(); }
    // term-weighted state (TermWeight IDF / PMI): token weights, float counts,
    // per-document weight sums
    bool weighted = false;
    std::vector<float> wt, ndk_f, nkw_f, nk_f, docw;
    uint32_t global_step = 0;
    uint32_t optim_interval = 10, burn_in = 0;
    // partition chunks from these word frequencies instead of the corpus's
    std::vector<uint64_t> chunk_cf;
};

void add_token(Model& m, int64_t d, uint32_t v, int k, int inc, int32_t* nk_local) {
    // addWordTo<+-1> (LDAModel.hpp:287-299) for TermWeight::one.
    m.ndk[(size_t)d * m.K + k] += inc;
    nk_local[k] += inc;
    m.nkw[(size_t)v * m.K + k] += inc;
}

// sampleDocument (LDAModel.hpp:326-354) over the token range [b, e).
void sample_range(Model& m, int64_t d, uint64_t b, uint64_t e, int32_t* nk_local, uint32_t sweep, float* buf) {
    for (uint64_t g = b; g < e; ++g) {
        uint32_t v = m.words[g];
        if (v >= (uint32_t)m.V) continue;
        int zo = m.z[g];
        add_token(m, d, v, zo, -1, nk_local);
        float u = token_uniform(m.seed, g, sweep, TAG_SWEEP);
        int zn = sample_one(m.K, m.scan_order, m.kpl, &m.ndk[(size_t)d * m.K], m.alpha.data(),
                            &m.nkw[(size_t)v * m.K], m.eta, nk_local, m.veta, u, buf, m.ecol(v), m.esum());
        m.z[g] = (uint16_t)zn;
        add_token(m, d, v, zn, +1, nk_local);
    }
}

// Vocabulary chunks of the partition scheme (LDAModel.hpp:436-449): boundaries
// at cumulative collection-frequency quantiles.  Missing trailing chunks (the
// reference can produce fewer than W) are padded with V (empty chunks).
std::vector<uint32_t> vocab_chunks(const Model& m, int W) {
    std::vector<uint64_t> cf(m.V, 0);
    if (!m.chunk_cf.empty()) cf = m.chunk_cf;  // the model's vocabCf (fold-in together)
    else for (uint32_t v : m.words) if (v < (uint32_t)m.V) ++cf[v];
    uint64_t tot = 0;
    for (auto c : cf) tot += c;
    std::vector<uint32_t> ch;
    uint64_t cum = 0;
    for (int i = 0; i < m.V; ++i) {
        cum += cf[i];
        if (cum * (uint64_t)W >= tot * (uint64_t)(ch.size() + 1)) ch.push_back((uint32_t)(i + 1));
    }
    while ((int)ch.size() < W) ch.push_back((uint32_t)m.V);
    ch.resize(W);
    ch[W - 1] = (uint32_t)m.V;
    return ch;
}

// One training sweep of the partition scheme with W workers
// (LDAModel.hpp:373-399, merge 571-579, distribute 596-602).  W == 1 is the
// `none` scheme (LDAModel.hpp:361-371).  Worker p keeps a private n_k that
// starts from the global n_k; round i gives worker p the documents
// d = id*W + ((i+p) % W) restricted to vocabulary chunk p; afterwards the
// global n_k is the row sum of n_kw.  Cells of a round are disjoint in
// documents and vocabulary, so running them on `threads` threads gives the
// same result as running them one by one.
void sweep_partition(Model& m, int W, int threads) {
    const uint32_t sweep = m.global_step;
    std::vector<uint32_t> ch = vocab_chunks(m, W);
    // chunk offsets inside each document (LDAModel.hpp:451-469)
    std::vector<uint64_t> coff((size_t)(W + 1) * m.D);
    for (int64_t d = 0; d < m.D; ++d) {
        uint64_t b = m.off[d], e = m.off[d + 1];
        coff[d] = b;
        for (int p = 0; p < W; ++p) {
            uint32_t lim = ch[p];
            coff[(size_t)(p + 1) * m.D + d] = (uint64_t)(std::lower_bound(m.words.begin() + b, m.words.begin() + e, lim) - m.words.begin());
        }
        coff[(size_t)W * m.D + d] = e;
    }
    std::vector<int32_t> nk_priv((size_t)W * m.K);
    for (int p = 0; p < W; ++p) std::copy(m.nk.begin(), m.nk.end(), nk_priv.begin() + (size_t)p * m.K);
    int T = std::max(1, std::min(threads, W));
    for (int i = 0; i < W; ++i) {
        auto work = [&](int tid) {
            std::vector<float> buf(64 * device_kpl(m.K) + m.K + 64);
            std::vector<uint32_t> order;
            for (int p = tid; p < W; p += T) {
                int didx = (i + p) % W;
                size_t n = (size_t)((m.D + (W - 1) - didx) / W);
                uint32_t sseed = philox_word(m.seed, (uint32_t)p, (uint32_t)i, sweep, TAG_SHUFFLE);
                shuffled_order(n, sseed, order);
                for (size_t q = 0; q < n; ++q) {
                    int64_t d = (int64_t)order[q] * W + didx;
                    sample_range(m, d, coff[(size_t)p * m.D + d], coff[(size_t)(p + 1) * m.D + d],
                                 &nk_priv[(size_t)p * m.K], sweep, buf.data());
                }
            }
        };
        if (T == 1) work(0);
        else {
            std::vector<std::thread> th;
            for (int t = 0; t < T; ++t) th.emplace_back(work, t);
            for (auto& t : th) t.join();
        }
    }
    // mergeState<partition>: n_k = rowwise sum of n_kw
    std::fill(m.nk.begin(), m.nk.end(), 0);
    for (int v = 0; v < m.V; ++v)
        for (int k = 0; k < m.K; ++k) m.nk[k] += m.nkw[(size_t)v * m.K + k];
}

// One training sweep of the copy_merge scheme with W workers
// (LDAModel.hpp:401-423, merge 555-570, distribute 586-595).  Every worker
// starts from a private copy of n_kw and n_k; the documents are cut into
// chStride = min(8W, D) strided tasks, task c visiting documents id*chStride + c
// in forShuffled order.  tomoto hands tasks to whichever thread is free; here
// task c belongs to worker c % W, which makes the schedule deterministic.
// Merge: n_kw = old + sum_w (local_w - old), then n_k = rowwise sum of n_kw.
void sweep_copy_merge(Model& m, int W, int threads) {
    const uint32_t sweep = m.global_step;
    const int64_t chStride = std::min<int64_t>((int64_t)W * 8, m.D);
    std::vector<std::vector<int32_t>> nkw_loc(W, m.nkw);
    std::vector<std::vector<int32_t>> nk_loc(W, m.nk);
    auto work = [&](int w) {
        std::vector<float> buf(64 * device_kpl(m.K) + m.K + 64);
        std::vector<uint32_t> order;
        for (int64_t ch = w; ch < chStride; ch += W) {
            size_t n = (size_t)((m.D + (chStride - 1) - ch) / chStride);
            uint32_t sseed = philox_word(m.seed, (uint32_t)ch, 0xC0u, sweep, TAG_SHUFFLE);
            shuffled_order(n, sseed, order);
            for (size_t q = 0; q < n; ++q) {
                int64_t d = (int64_t)order[q] * chStride + ch;
                for (uint64_t g = m.off[d]; g < m.off[d + 1]; ++g) {
                    uint32_t v = m.words[g];
                    if (v >= (uint32_t)m.V) continue;
                    int32_t* ndk = &m.ndk[(size_t)d * m.K];
                    int32_t* col = &nkw_loc[w][(size_t)v * m.K];
                    int32_t* nk = nk_loc[w].data();
                    int zo = m.z[g];
                    ndk[zo]--; col[zo]--; nk[zo]--;
                    float u = token_uniform(m.seed, g, sweep, TAG_SWEEP);
                    int zn = sample_one(m.K, m.scan_order, m.kpl, ndk, m.alpha.data(), col, m.eta, nk, m.veta, u, buf.data(),
                                        m.ecol(v), m.esum());
                    m.z[g] = (uint16_t)zn;
                    ndk[zn]++; col[zn]++; nk[zn]++;
                }
            }
        }
    };
    int T = std::max(1, std::min(threads, W));
    if (T == 1) for (int w = 0; w < W; ++w) work(w);
    else {
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t) th.emplace_back([&, t] { for (int w = t; w < W; w += T) work(w); });
        for (auto& t : th) t.join();
    }
    std::vector<int32_t> old = m.nkw;
    for (size_t i = 0; i < m.nkw.size(); ++i) {
        int32_t acc = old[i];
        for (int w = 0; w < W; ++w) acc += nkw_loc[w][i] - old[i];
        m.nkw[i] = acc;
    }
    std::fill(m.nk.begin(), m.nk.end(), 0);
    for (int v = 0; v < m.V; ++v)
        for (int k = 0; k < m.K; ++k) m.nk[k] += m.nkw[(size_t)v * m.K + k];
}

// optimizeParameters (LDAModel.hpp:221-265): 10 Minka fixed-point passes.
// Terms are float digammaApprox as in tomoto; sums are accumulated in double
// (tomoto sums floats in Eigen's order); dAlpha is tomoto's LUT digamma.
void optimize_alpha(Model& m) {
    std::vector<int32_t> len(m.D);
    for (int64_t d = 0; d < m.D; ++d) {
        int32_t c = 0;
        for (uint64_t g = m.off[d]; g < m.off[d + 1]; ++g) c += m.words[g] < (uint32_t)m.V;
        len[d] = c;
    }
    for (int it = 0; it < 10; ++it) {
        float asum = 0.0f;
        for (int k = 0; k < m.K; ++k) asum += m.alpha[k];
        double s = 0.0;
        for (int64_t d = 0; d < m.D; ++d) s += (double)digamma_approx_f((float)len[d] + asum);
        float denom = (float)(s - (double)m.D * (double)lut_digamma().get(asum));
        for (int k = 0; k < m.K; ++k) {
            float ak = m.alpha[k];
            double sk = 0.0;
            for (int64_t d = 0; d < m.D; ++d) sk += (double)digamma_approx_f((float)m.ndk[(size_t)d * m.K + k] + ak);
            float nom = (float)(sk - (double)m.D * (double)lut_digamma().get(ak));
            m.alpha[k] = std::max(nom / denom * ak, 1e-5f);
        }
    }
}

// getLLDocs / getLLRest (LDAModel.hpp:616-674) for symmetric eta.
// mode 0: FP64 (double lgamma, double sums) -- the north-star LL.
// mode 1: tomoto (float LUT lgamma terms, double accumulation).
void loglik(const Model& m, int mode, double* ll_docs, double* ll_rest) {
    auto lg = [&](double x) -> double { return mode ? (double)lut_lgamma().get((float)x) : std::lgamma(x); };
    double asum_d = 0.0; float asum_f = 0.0f;
    for (int k = 0; k < m.K; ++k) { asum_d += (double)m.alpha[k]; asum_f += m.alpha[k]; }
    double asum = mode ? (double)asum_f : asum_d;
    double docs = 0.0;
    for (int64_t d = 0; d < m.D; ++d) {
        int64_t L = 0;
        for (uint64_t g = m.off[d]; g < m.off[d + 1]; ++g) L += m.words[g] < (uint32_t)m.V;
        double Ld = mode ? (double)((float)L + asum_f) : (double)L + asum;
        docs -= lg(Ld) - lg(asum);
        for (int k = 0; k < m.K; ++k) {
            int32_t c = m.ndk[(size_t)d * m.K + k];
            double x = mode ? (double)((float)c + m.alpha[k]) : (double)c + (double)m.alpha[k];
            docs += lg(x) - lg((double)m.alpha[k]);
        }
    }
    double eta = (double)m.eta;
    double veta = mode ? (double)m.veta : (double)m.V * eta;
    double rest = 0.0;
    double lge = lg(eta);
    if (!m.eta_kv.empty()) {
        // word priors (LDAModel.hpp:638-650): per-topic eta sums and eta_kv
        for (int k = 0; k < m.K; ++k) {
            double es = mode ? (double)m.etasum[k] : m.etasum_d[k];
            double x = mode ? (double)((float)m.nk[k] + m.etasum[k]) : (double)m.nk[k] + es;
            rest += lg(es) - lg(x);
            for (int v = 0; v < m.V; ++v) {
                int32_t c = m.nkw[(size_t)v * m.K + k];
                if (!c) continue;
                float e = m.eta_kv[(size_t)v * m.K + k];
                double y = mode ? (double)((float)c + e) : (double)c + (double)e;
                rest += lg(y) - lg((double)e);
            }
        }
        *ll_docs = docs;
        *ll_rest = rest;
        return;
    }
    rest = lg(veta) * m.K;
    for (int k = 0; k < m.K; ++k) {
        double x = mode ? (double)((float)m.nk[k] + m.veta) : (double)m.nk[k] + veta;
        rest -= lg(x);
        for (int v = 0; v < m.V; ++v) {
            int32_t c = m.nkw[(size_t)v * m.K + k];
            if (!c) continue;
            double y = mode ? (double)((float)c + m.eta) : (double)c + eta;
            rest += lg(y) - lge;
        }
    }
    *ll_docs = docs;
    *ll_rest = rest;
}

// ------------------------------------------------ term-weighted (IDF/PMI) --
// TermWeight IDF / PMI (LDAModel.hpp:287-299, Utils.hpp:132-155): every token
// carries a float weight w; adding a token is val += w, removing it is
// val = max(val - w, 0) (std::max keeps val - w unless it is negative), and
// n_dk, n_kw, n_k are float.
inline float sub_clamp(float v, float w) {
    float a = v - w;
    return a < 0.0f ? 0.0f : a;
}

inline void add_token_w(Model& m, int64_t d, uint32_t v, int k, float w, bool dec, float* nk_local) {
    float& a = m.ndk_f[(size_t)d * m.K + k];
    float& b = nk_local[k];
    float& c = m.nkw_f[(size_t)v * m.K + k];
    if (dec) { a = sub_clamp(a, w); b = sub_clamp(b, w); c = sub_clamp(c, w); }
    else { a += w; b += w; c += w; }
}

// getZLikelihoods (LDAModel.hpp:273-285) over float counts, then the same
// prefix sum and search as the integer sampler.
int sample_one_w(const Model& m, const float* ndk, const float* col, const float* nk, float u, float* buf,
                 const float* ecol) {
    int n = m.scan_order ? 64 * m.kpl : m.K;
    const float* ec = ecol;
    const float* es = m.esum();
    for (int k = 0; k < m.K; ++k) {
        float a = (ndk[k] + m.alpha[k]) * (col[k] + (ec ? ec[k] : m.eta));
        buf[k] = a / (nk[k] + (es ? es[k] : m.veta));
    }
    for (int k = m.K; k < n; ++k) buf[k] = 0.0f;
    if (m.scan_order) prefix_sum_device(buf, m.kpl);
    else prefix_sum_tomoto(buf, m.K);
    return search_cdf(buf, m.K, u);
}

void sample_range_w(Model& m, int64_t d, uint64_t b, uint64_t e, float* nk_local, uint32_t sweep, float* buf) {
    for (uint64_t g = b; g < e; ++g) {
        uint32_t v = m.words[g];
        if (v >= (uint32_t)m.V) continue;
        const float w = m.wt[g];
        int zo = m.z[g];
        add_token_w(m, d, v, zo, w, true, nk_local);
        float u = token_uniform(m.seed, g, sweep, TAG_SWEEP);
        int zn = sample_one_w(m, &m.ndk_f[(size_t)d * m.K], &m.nkw_f[(size_t)v * m.K], nk_local, u, buf, m.ecol(v));
        m.z[g] = (uint16_t)zn;
        add_token_w(m, d, v, zn, w, false, nk_local);
    }
}

// mergeState<partition> for float counts (LDAModel.hpp:572-579): n_kw =
// max(n_kw, 0) cell by cell, then n_k = the row sums of n_kw in float, in a
// fixed two-level order: words in blocks of 256, each block summed in
// ascending word order, then the block sums in ascending block order.
// (Eigen's rowwise().sum() groups the terms by its SIMD packet width, so
// tomoto's own order depends on the build; the device follows this one.)
constexpr int kRowsumBlock = 256;
void merge_w(Model& m) {
    for (auto& x : m.nkw_f) if (x < 0.0f) x = 0.0f;
    std::fill(m.nk_f.begin(), m.nk_f.end(), 0.0f);
    std::vector<float> part(m.K);
    for (int v0 = 0; v0 < m.V; v0 += kRowsumBlock) {
        std::fill(part.begin(), part.end(), 0.0f);
        for (int v = v0; v < std::min(m.V, v0 + kRowsumBlock); ++v)
            for (int k = 0; k < m.K; ++k) part[k] = part[k] + m.nkw_f[(size_t)v * m.K + k];
        for (int k = 0; k < m.K; ++k) m.nk_f[k] = m.nk_f[k] + part[k];
    }
}

// Partition sweep (the schedule of sweep_partition) with token weights.
void sweep_partition_w(Model& m, int W, int threads) {
    const uint32_t sweep = m.global_step;
    std::vector<uint32_t> ch = vocab_chunks(m, W);
    std::vector<uint64_t> coff((size_t)(W + 1) * m.D);
    for (int64_t d = 0; d < m.D; ++d) {
        uint64_t b = m.off[d], e = m.off[d + 1];
        coff[d] = b;
        for (int p = 0; p < W; ++p)
            coff[(size_t)(p + 1) * m.D + d] =
                (uint64_t)(std::lower_bound(m.words.begin() + b, m.words.begin() + e, ch[p]) - m.words.begin());
        coff[(size_t)W * m.D + d] = e;
    }
    std::vector<float> nk_priv((size_t)W * m.K);
    for (int p = 0; p < W; ++p) std::copy(m.nk_f.begin(), m.nk_f.end(), nk_priv.begin() + (size_t)p * m.K);
    int T = std::max(1, std::min(threads, W));
    for (int i = 0; i < W; ++i) {
        auto work = [&](int tid) {
            std::vector<float> buf(64 * device_kpl(m.K) + m.K + 64);
            std::vector<uint32_t> order;
            for (int p = tid; p < W; p += T) {
                int didx = (i + p) % W;
                size_t n = (size_t)((m.D + (W - 1) - didx) / W);
                uint32_t sseed = philox_word(m.seed, (uint32_t)p, (uint32_t)i, sweep, TAG_SHUFFLE);
                shuffled_order(n, sseed, order);
                for (size_t q = 0; q < n; ++q) {
                    int64_t d = (int64_t)order[q] * W + didx;
                    sample_range_w(m, d, coff[(size_t)p * m.D + d], coff[(size_t)(p + 1) * m.D + d],
                                   &nk_priv[(size_t)p * m.K], sweep, buf.data());
                }
            }
        };
        if (T == 1) work(0);
        else {
            std::vector<std::thread> th;
            for (int t = 0; t < T; ++t) th.emplace_back(work, t);
            for (auto& t : th) t.join();
        }
    }
    merge_w(m);
}

// copy_merge with token weights (the schedule of sweep_copy_merge): worker
// w samples on private copies of n_kw and n_k; the merge follows
// mergeState<copy_merge> (LDAModel.hpp:555-569) in float: n_kw = local_0, then
// n_kw += local_i - old for i = 1 .. W-1 in order, each difference rounded
// first; then the weighted clamp and row sum of merge_w.
void sweep_copy_merge_w(Model& m, int W, int threads) {
    const uint32_t sweep = m.global_step;
    const int64_t chStride = std::min<int64_t>((int64_t)W * 8, m.D);
    std::vector<std::vector<float>> nkw_loc(W, m.nkw_f);
    std::vector<std::vector<float>> nk_loc(W, m.nk_f);
    auto work = [&](int w) {
        std::vector<float> buf(64 * device_kpl(m.K) + m.K + 64);
        std::vector<uint32_t> order;
        for (int64_t ch = w; ch < chStride; ch += W) {
            size_t n = (size_t)((m.D + (chStride - 1) - ch) / chStride);
            uint32_t sseed = philox_word(m.seed, (uint32_t)ch, 0xC0u, sweep, TAG_SHUFFLE);
            shuffled_order(n, sseed, order);
            for (size_t q = 0; q < n; ++q) {
                int64_t d = (int64_t)order[q] * chStride + ch;
                for (uint64_t g = m.off[d]; g < m.off[d + 1]; ++g) {
                    uint32_t v = m.words[g];
                    if (v >= (uint32_t)m.V) continue;
                    const float wt = m.wt[g];
                    float* ndk = &m.ndk_f[(size_t)d * m.K];
                    float* col = &nkw_loc[w][(size_t)v * m.K];
                    float* nk = nk_loc[w].data();
                    int zo = m.z[g];
                    ndk[zo] = sub_clamp(ndk[zo], wt); col[zo] = sub_clamp(col[zo], wt); nk[zo] = sub_clamp(nk[zo], wt);
                    float u = token_uniform(m.seed, g, sweep, TAG_SWEEP);
                    int zn = sample_one_w(m, ndk, col, nk, u, buf.data(), m.ecol(v));
                    m.z[g] = (uint16_t)zn;
                    ndk[zn] += wt; col[zn] += wt; nk[zn] += wt;
                }
            }
        }
    };
    int T = std::max(1, std::min(threads, W));
    if (T == 1) for (int w = 0; w < W; ++w) work(w);
    else {
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t) th.emplace_back([&, t] { for (int w = t; w < W; w += T) work(w); });
        for (auto& t : th) t.join();
    }
    for (size_t i = 0; i < m.nkw_f.size(); ++i) {
        const float old = m.nkw_f[i];
        float acc = nkw_loc[0][i];
        for (int w = 1; w < W; ++w) acc = acc + (nkw_loc[w][i] - old);
        m.nkw_f[i] = acc;
    }
    merge_w(m);
}

// optimizeParameters (LDAModel.hpp:221-265) with the document length
// replaced by the document's weight sum (getSumWordWeight).
void optimize_alpha_w(Model& m) {
    for (int it = 0; it < 10; ++it) {
        float asum = 0.0f;
        for (int k = 0; k < m.K; ++k) asum += m.alpha[k];
        double s = 0.0;
        for (int64_t d = 0; d < m.D; ++d) s += (double)digamma_approx_f(m.docw[d] + asum);
        float denom = (float)(s - (double)m.D * (double)lut_digamma().get(asum));
        for (int k = 0; k < m.K; ++k) {
            float ak = m.alpha[k];
            double sk = 0.0;
            for (int64_t d = 0; d < m.D; ++d) sk += (double)digamma_approx_f(m.ndk_f[(size_t)d * m.K + k] + ak);
            float nom = (float)(sk - (double)m.D * (double)lut_digamma().get(ak));
            m.alpha[k] = std::max(nom / denom * ak, 1e-5f);
        }
    }
}

// getLLDocs + getLLRest (LDAModel.hpp:616-666) in FP64 over float counts;
// the document term uses the weight sum.  Zero cells add lgamma(x) - lgamma(x)
// and are skipped.
void loglik_w(const Model& m, double* ll_docs, double* ll_rest) {
    double asum = 0.0;
    for (int k = 0; k < m.K; ++k) asum += (double)m.alpha[k];
    double docs = 0.0;
    for (int64_t d = 0; d < m.D; ++d) {
        docs -= std::lgamma((double)m.docw[d] + asum) - std::lgamma(asum);
        for (int k = 0; k < m.K; ++k) {
            float c = m.ndk_f[(size_t)d * m.K + k];
            if (c == 0.0f) continue;
            docs += std::lgamma((double)c + (double)m.alpha[k]) - std::lgamma((double)m.alpha[k]);
        }
    }
    double eta = (double)m.eta, veta = (double)m.V * eta;
    double rest = 0.0, lge = std::lgamma(eta);
    for (int k = 0; k < m.K; ++k) {
        double es = m.eta_kv.empty() ? veta : m.etasum_d[k];
        rest += std::lgamma(es) - std::lgamma((double)m.nk_f[k] + es);
        for (int v = 0; v < m.V; ++v) {
            float c = m.nkw_f[(size_t)v * m.K + k];
            if (c == 0.0f) continue;
            if (m.eta_kv.empty()) rest += std::lgamma((double)c + eta) - lge;
            else {
                double e = (double)m.eta_kv[(size_t)v * m.K + k];
                rest += std::lgamma((double)c + e) - std::lgamma(e);
            }
        }
    }
    *ll_docs = docs;
    *ll_rest = rest;
}

}  // namespace

// ============================================================== C API ======
extern "C" {

void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { philox4x32_10(ctr, key, out); }

float orc_token_uniform(uint64_t seed, uint64_t g, uint32_t sweep, uint32_t tag) { return token_uniform(seed, g, sweep, tag); }

uint32_t orc_tag(int which) {
    const uint32_t t[5] = {TAG_SWEEP, TAG_INIT, TAG_SHUFFLE, TAG_INFER, TAG_INFINIT};
    return (which >= 0 && which < 5) ? t[which] : 0u;
}

int orc_device_kpl(int K) { return device_kpl(K); }

void orc_weights(int K, const int32_t* ndk, const float* alpha, const int32_t* nkw, float eta,
                 const int32_t* nk, float veta, float* out) {
    for (int k = 0; k < K; ++k) out[k] = topic_weight(ndk[k], alpha[k], nkw[k], eta, nk[k], veta);
}

void orc_prefix_sum_tomoto(float* x, int n) { prefix_sum_tomoto(x, n); }
void orc_prefix_sum_device(float* x, int kpl) { prefix_sum_device(x, kpl); }
int orc_search(const float* cdf, int K, float u) { return search_cdf(cdf, K, u); }

void orc_shuffled_order(uint64_t N, uint32_t seed, uint32_t* out) {
    std::vector<uint32_t> o;
    shuffled_order((size_t)N, seed, o);
    std::copy(o.begin(), o.end(), out);
}

// Markstein-corrected quotient with a correctly rounded reciprocal:
// q0 = a*y, r = fma(-q0, b, a), q = fma(r, y, q0).  Returns the number of
// (a, b) pairs among n pseudo-random ones (a in [amin, amax), b in
// [bmin, bmax), log-uniform) where q differs from the IEEE quotient a/b.
uint64_t orc_markstein_mismatches(uint64_t n, uint64_t seed, double amin, double amax, double bmin, double bmax) {
    uint64_t bad = 0;
    double la = std::log(amin), ra = std::log(amax) - la, lb = std::log(bmin), rb = std::log(bmax) - lb;
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t ctr[4] = {(uint32_t)i, (uint32_t)(i >> 32), 7u, 0x4D41524Bu};
        uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
        uint32_t o[4];
        philox4x32_10(ctr, key, o);
        float a = (float)std::exp(la + ra * (o[0] * 0x1p-32));
        float b = (float)std::exp(lb + rb * (o[1] * 0x1p-32));
        if (o[2] & 1) b = std::nearbyint(b) + 0.01f * 5000.0f;  // integer count + V*eta shapes
        float y = 1.0f / b;
        float q0 = a * y;
        float r = std::fma(-q0, b, a);
        float q = std::fma(r, y, q0);
        float ref = a / b;
        if (std::memcmp(&q, &ref, 4) != 0) ++bad;
    }
    return bad;
}

// Frozen-state sampling (the pure per-token function): for each query token
// t, remove its own assignment z_old[t] from (ndk row, nkw column, nk), then
// compute weights, scan and search with uniform u[t].  ndk is D x K, nkw is
// V x K (word-major), nk is K.  Counts are not modified.
void orc_sample_frozen(int K, int scan_order, int kpl, const float* alpha, float eta, float veta,
                       const int32_t* ndk, const int32_t* nkw, const int32_t* nk,
                       int64_t n, const int64_t* doc, const uint32_t* word, const uint16_t* z_old,
                       const float* u, uint16_t* z_new, float* total_out) {
    std::vector<int32_t> a(K), c(K), e(K);
    std::vector<float> buf(64 * device_kpl(K) + K + 64);
    for (int64_t t = 0; t < n; ++t) {
        std::copy(ndk + (size_t)doc[t] * K, ndk + (size_t)doc[t] * K + K, a.begin());
        std::copy(nkw + (size_t)word[t] * K, nkw + (size_t)word[t] * K + K, c.begin());
        std::copy(nk, nk + K, e.begin());
        int zo = z_old[t];
        a[zo] -= 1; c[zo] -= 1; e[zo] -= 1;
        z_new[t] = (uint16_t)sample_one(K, scan_order, kpl, a.data(), alpha, c.data(), eta, e.data(), veta, u[t], buf.data());
        if (total_out) total_out[t] = buf[K - 1];
    }
}

float orc_lgamma_lut(float x) { return lut_lgamma().get(x); }
float orc_digamma_lut(float x) { return lut_digamma().get(x); }
float orc_digamma_approx(float x) { return digamma_approx_f(x); }

// ---- model handle ----
void* orc_create(int K, int V, const float* alpha, float eta, uint64_t seed, int scan_order) {
    Model* m = new Model();
    m->K = K; m->V = V; m->eta = eta; m->seed = seed;
    m->veta = eta * (float)V;
    m->kpl = device_kpl(K);
    m->scan_order = scan_order;
    m->alpha.assign(alpha, alpha + K);
    m->nkw.assign((size_t)V * K, 0);
    m->nk.assign(K, 0);
    return m;
}

void orc_destroy(void* h) { delete (Model*)h; }

void orc_set_optim(void* h, uint32_t optim_interval, uint32_t burn_in) {
    Model* m = (Model*)h;
    m->optim_interval = optim_interval; m->burn_in = burn_in;
}

// words: sorted ascending within each document.  z may be null (then the
// initial topics are drawn as below).
void orc_set_corpus(void* h, const uint32_t* words, const uint64_t* off, int64_t D, const uint16_t* z) {
    Model* m = (Model*)h;
    m->D = D;
    m->off.assign(off, off + D + 1);
    uint64_t N = off[D];
    m->words.assign(words, words + N);
    m->z.assign(N, 0xFFFF);
    m->ndk.assign((size_t)D * m->K, 0);
    std::fill(m->nkw.begin(), m->nkw.end(), 0);
    std::fill(m->nk.begin(), m->nk.end(), 0);
    // initial draw z ~ Discrete(alpha) (LDAModel.hpp:754-762, 776-813): first k
    // with r < cdf_alpha[k], cdf_alpha the sequential float sum of alpha.
    std::vector<float> ca(m->K);
    float s = 0.0f;
    for (int k = 0; k < m->K; ++k) { s = s + m->alpha[k]; ca[k] = s; }
    for (int64_t d = 0; d < D; ++d) {
        for (uint64_t g = off[d]; g < off[d + 1]; ++g) {
            uint32_t v = words[g];
            if (v >= (uint32_t)m->V) continue;
            int k = z ? z[g]
                : m->ecol(v) ? sample_from_discrete(m->ecol(v), m->K, token_uniform(m->seed, g, 0, TAG_INIT))
                             : search_cdf(ca.data(), m->K, token_uniform(m->seed, g, 0, TAG_INIT));
            m->z[g] = (uint16_t)k;
            add_token(*m, d, v, k, +1, m->nk.data());
        }
    }
}

// Word priors: eta_kv (V x K, natural order); tomoto builds it as eta
// everywhere, replaced by the given columns (prepareWordPriors,
// LDAModel.hpp:724-737).  Call before orc_set_corpus (the initial draw uses it).
void orc_set_priors(void* h, const float* eta_kv) {
    Model* m = (Model*)h;
    m->eta_kv.assign(eta_kv, eta_kv + (size_t)m->V * m->K);
    m->etasum.assign(m->K, 0.0f);
    m->etasum_d.assign(m->K, 0.0);
    for (int v = 0; v < m->V; ++v)
        for (int k = 0; k < m->K; ++k) {
            m->etasum[k] = m->etasum[k] + eta_kv[(size_t)v * m->K + k];
            m->etasum_d[k] += (double)eta_kv[(size_t)v * m->K + k];
        }
}

int orc_sample_from_discrete(const float* w, int K, float u) { return sample_from_discrete(w, K, u); }

void orc_set_alpha(void* h, const float* alpha) { Model* m = (Model*)h; m->alpha.assign(alpha, alpha + m->K); }
void orc_get_alpha(void* h, float* alpha) { Model* m = (Model*)h; std::copy(m->alpha.begin(), m->alpha.end(), alpha); }
uint32_t orc_global_step(void* h) { return ((Model*)h)->global_step; }

void orc_train_scheme(void* h, int iterations, int workers, int threads, int scheme);

// TopicModel::train (TopicModel.hpp:684-755) + trainOne (LDAModel.hpp:505-539):
// per iteration one sweep (partition with W workers; W == 1 is `none`),
// then alpha optimisation when globalStep >= burnIn and
// (globalStep+1) % optimInterval == 0, then ++globalStep.
void orc_train(void* h, int iterations, int workers, int threads) { orc_train_scheme(h, iterations, workers, threads, 3); }

// scheme: 1 none, 2 copy_merge, 3 partition (tomotopy.ParallelScheme values)
void orc_train_scheme(void* h, int iterations, int workers, int threads, int scheme) {
    Model* m = (Model*)h;
    int W = std::max(1, workers);
    if (scheme == 1) W = 1;
    for (int it = 0; it < iterations; ++it) {
        if (scheme == 2 && W > 1) sweep_copy_merge(*m, W, threads);
        else sweep_partition(*m, W, threads);
        if (m->global_step >= m->burn_in && m->optim_interval && (m->global_step + 1) % m->optim_interval == 0)
            optimize_alpha(*m);
        ++m->global_step;
    }
}

void orc_optimize_alpha(void* h) { optimize_alpha(*(Model*)h); }

void orc_get_state(void* h, uint16_t* z, int32_t* ndk, int32_t* nkw, int32_t* nk) {
    Model* m = (Model*)h;
    if (z) std::copy(m->z.begin(), m->z.end(), z);
    if (ndk) std::copy(m->ndk.begin(), m->ndk.end(), ndk);
    if (nkw) std::copy(m->nkw.begin(), m->nkw.end(), nkw);
    if (nk) std::copy(m->nk.begin(), m->nk.end(), nk);
}

void orc_loglik(void* h, int mode, double* ll_docs, double* ll_rest) { loglik(*(Model*)h, mode, ll_docs, ll_rest); }

// Fold-in with together=True (LDAModel.hpp:852-892): the unseen documents
// join a copy of the model's n_kw / n_k (their initial topics drawn from
// Discrete(alpha), keyed by infer_seed), then `iters` sweeps of the scheme
// (1 none, 2 copy_merge, 3 partition with W workers; partition chunks from the
// model's word frequencies, as updatePartition uses vocabCf) run over them
// with alpha fixed; ll = getLLRest(copy) - getLLRest(model) + getLLDocs(docs)
// in FP64, theta = (n_dk + alpha) / (L + sum alpha).
void orc_infer_together(void* h, const uint32_t* words, const uint64_t* off, int64_t D, int iters, uint64_t infer_seed,
                        int workers, int scheme, int threads, float* theta, double* ll, uint16_t* z_out) {
    const Model* m = (const Model*)h;
    Model t;
    t.K = m->K; t.V = m->V; t.kpl = m->kpl; t.scan_order = m->scan_order;
    t.eta = m->eta; t.veta = m->veta; t.seed = infer_seed;
    t.alpha = m->alpha;
    t.nkw.assign((size_t)m->V * m->K, 0);
    t.nk.assign(m->K, 0);
    t.chunk_cf.assign(m->V, 0);
    for (uint32_t v : m->words) if (v < (uint32_t)m->V) ++t.chunk_cf[v];
    t.optim_interval = 0;
    // word priors: the copy keeps eta_kv (initial topics from the eta columns)
    t.eta_kv = m->eta_kv; t.etasum = m->etasum; t.etasum_d = m->etasum_d;
    orc_set_corpus(&t, words, off, D, nullptr);
    for (size_t i = 0; i < t.nkw.size(); ++i) t.nkw[i] += m->nkw[i];
    for (int k = 0; k < m->K; ++k) t.nk[k] += m->nk[k];
    orc_train_scheme(&t, iters, workers, threads, scheme);
    double docs = 0, rest = 0, docs0 = 0, rest0 = 0;
    loglik(t, 0, &docs, &rest);
    loglik(*m, 0, &docs0, &rest0);
    *ll = rest - rest0 + docs;
    if (z_out) std::copy(t.z.begin(), t.z.end(), z_out);
    float asum = 0.0f;
    for (int k = 0; k < m->K; ++k) asum += m->alpha[k];
    for (int64_t d = 0; d < D; ++d) {
        int32_t L = 0;
        for (int k = 0; k < m->K; ++k) L += t.ndk[(size_t)d * m->K + k];
        for (int k = 0; k < m->K; ++k)
            theta[(size_t)d * m->K + k] = ((float)t.ndk[(size_t)d * m->K + k] + m->alpha[k]) / ((float)L + asum);
    }
}

// Term-weighted corpus (TermWeight IDF / PMI): weights[g] is token g's weight
// (ignored for ids >= V).  Initial topics as in orc_set_corpus; the float
// counts are accumulated token by token in corpus order (initializeDocState,
// LDAModel.hpp:782-813), and each document's weight sum is the sequential
// float sum of its tokens' weights (LDA.h:113-116).
void orc_set_corpus_w(void* h, const uint32_t* words, const uint64_t* off, int64_t D, const uint16_t* z,
                      const float* weights) {
    Model* m = (Model*)h;
    orc_set_corpus(h, words, off, D, z);
    const uint64_t N = off[D];
    m->weighted = true;
    m->wt.assign(weights, weights + N);
    m->ndk_f.assign((size_t)D * m->K, 0.0f);
    m->nkw_f.assign((size_t)m->V * m->K, 0.0f);
    m->nk_f.assign(m->K, 0.0f);
    m->docw.assign(D, 0.0f);
    for (int64_t d = 0; d < D; ++d) {
        float sw = 0.0f;
        for (uint64_t g = off[d]; g < off[d + 1]; ++g) {
            uint32_t v = words[g];
            if (v >= (uint32_t)m->V) continue;
            add_token_w(*m, d, v, m->z[g], m->wt[g], false, m->nk_f.data());
            sw += m->wt[g];
        }
        m->docw[d] = sw;
    }
}

// Weighted training: scheme 1 (none), 2 (copy_merge) or 3 (partition, W
// workers), alpha optimisation on the same schedule as orc_train_scheme.
void orc_train_w(void* h, int iterations, int workers, int threads, int scheme) {
    Model* m = (Model*)h;
    int W = scheme == 1 ? 1 : std::max(1, workers);
    for (int it = 0; it < iterations; ++it) {
        if (scheme == 2 && W > 1) sweep_copy_merge_w(*m, W, threads);
        else sweep_partition_w(*m, W, threads);
        if (m->global_step >= m->burn_in && m->optim_interval && (m->global_step + 1) % m->optim_interval == 0)
            optimize_alpha_w(*m);
        ++m->global_step;
    }
}

void orc_optimize_alpha_w(void* h) { optimize_alpha_w(*(Model*)h); }
void orc_merge_w(void* h) { merge_w(*(Model*)h); }

void orc_get_state_w(void* h, uint16_t* z, float* ndk, float* nkw, float* nk, float* docw) {
    Model* m = (Model*)h;
    if (z) std::copy(m->z.begin(), m->z.end(), z);
    if (ndk) std::copy(m->ndk_f.begin(), m->ndk_f.end(), ndk);
    if (nkw) std::copy(m->nkw_f.begin(), m->nkw_f.end(), nkw);
    if (nk) std::copy(m->nk_f.begin(), m->nk_f.end(), nk);
    if (docw) std::copy(m->docw.begin(), m->docw.end(), docw);
}

// Replace the float counts (e.g. with a state downloaded from the device).
void orc_set_state_w(void* h, const uint16_t* z, const float* ndk, const float* nkw, const float* nk) {
    Model* m = (Model*)h;
    if (z) std::copy(z, z + m->z.size(), m->z.begin());
    if (ndk) std::copy(ndk, ndk + m->ndk_f.size(), m->ndk_f.begin());
    if (nkw) std::copy(nkw, nkw + m->nkw_f.size(), m->nkw_f.begin());
    if (nk) std::copy(nk, nk + m->nk_f.size(), m->nk_f.begin());
}

void orc_loglik_w(void* h, double* ll_docs, double* ll_rest) { loglik_w(*(Model*)h, ll_docs, ll_rest); }

// Fold-in of unseen documents, together=False (LDAModel.hpp:920-953): every
// document starts its own RNG stream (tomoto default-constructs `rgc{}` per
// document, so streams do not depend on the model seed; here they are keyed
// by `infer_seed` and the token position inside the document), draws its
// initial topics from Discrete(alpha), adds its tokens into a private copy of
// n_kw / n_k, and runs `iters` sweeps.  The private copy is represented
// exactly by its delta: n_k,eff = n_k + n_dk and n_kw,eff = n_kw + c_d.
// ll = getLLRest(private) - getLLRest(global) + getLLDocs(doc), evaluated in
// FP64 (mode 0) or with tomoto's LUT (mode 1) over the touched terms only.
// theta = (n_dk + alpha) / (L + sum(alpha)) (LDAModel.hpp:1145-1159).
// Segmented fold-in (the device's rule for long documents): S = 1 when
// seg_tokens == 0 or n <= seg_tokens, else the smallest power of two >=
// ceil(n / seg_tokens), at most seg_max; segment w starts at the first
// word-run start at or after floor(n*w/S).  Each iteration every segment
// samples its tokens in order against n_dk as it was when the iteration
// started plus its own changes (n_k,eff likewise); the segments' changes are
// then summed.  With sync_tokens > 0 the iteration is cut into steps: in step
// j every segment samples its tokens [start + j*sync_tokens, start +
// (j+1)*sync_tokens) against n_dk at the start of the step plus its own
// changes, and the changes are summed after every step (the integer device
// kernel merges after every 64-token batch).  c_d needs no merging: a word's
// tokens lie in one segment.  S == 1 is the sequential sampler above.
static int seg_count(uint64_t n, uint32_t seg_tokens, int seg_max) {
    if (seg_tokens == 0 || n <= seg_tokens) return 1;
    uint64_t need = (n + seg_tokens - 1) / seg_tokens;
    int S = 1;
    while ((uint64_t)S < need && S < seg_max) S <<= 1;
    return S;
}

static uint64_t seg_start(const uint32_t* w, uint64_t n, int i, int S) {
    if (i <= 0) return 0;
    if (i >= S) return n;
    uint64_t x = n * (uint64_t)i / (uint64_t)S;
    if (x == 0) return 0;
    while (x < n && w[x] == w[x - 1]) ++x;
    return x;
}

void orc_infer_seg(void* h, const uint32_t* words, const uint64_t* off, int64_t D, int iters, uint64_t infer_seed,
                   int mode, uint32_t seg_tokens, int seg_max, uint32_t sync_tokens, float* theta, double* ll,
                   uint16_t* z_out);

// Segment count and starts of one document (S + 1 entries in starts).
int orc_seg_bounds(const uint32_t* words, uint64_t n, uint32_t seg_tokens, int seg_max, uint64_t* starts) {
    const int S = seg_count(n, seg_tokens, seg_max);
    for (int i = 0; i <= S; ++i) starts[i] = seg_start(words, n, i, S);
    return S;
}

void orc_infer(void* h, const uint32_t* words, const uint64_t* off, int64_t D, int iters,
               uint64_t infer_seed, int mode, float* theta, double* ll, uint16_t* z_out) {
    orc_infer_seg(h, words, off, D, iters, infer_seed, mode, 0, 1, 0, theta, ll, z_out);
}

void orc_infer_seg(void* h, const uint32_t* words, const uint64_t* off, int64_t D, int iters, uint64_t infer_seed,
                   int mode, uint32_t seg_tokens, int seg_max, uint32_t sync_tokens, float* theta, double* ll,
                   uint16_t* z_out) {
    Model* m = (Model*)h;
    const int K = m->K;
    std::vector<float> ca(K);
    float s = 0.0f;
    for (int k = 0; k < K; ++k) { s = s + m->alpha[k]; ca[k] = s; }
    float asum_f = s;
    double asum_d = 0.0;
    for (int k = 0; k < K; ++k) asum_d += (double)m->alpha[k];
    std::vector<float> buf(64 * device_kpl(K) + K + 64);
    auto lg = [&](double x) -> double { return mode ? (double)lut_lgamma().get((float)x) : std::lgamma(x); };
    for (int64_t d = 0; d < D; ++d) {
        uint64_t b = off[d], e = off[d + 1];
        uint64_t n = e - b;
        std::vector<uint16_t> z(n, 0xFFFF);
        std::vector<int32_t> ndk(K, 0), nk_eff(m->nk);
        // doc-local deltas of n_kw, keyed by (word, topic); words are sorted,
        // so a word's tokens are one contiguous run.
        std::vector<int32_t> cd((size_t)n * K, 0);  // run-start index x K
        std::vector<uint64_t> run_start(n);
        for (uint64_t i = 0; i < n; ++i) run_start[i] = (i && words[b + i] == words[b + i - 1]) ? run_start[i - 1] : i;
        for (uint64_t i = 0; i < n; ++i) {
            uint32_t v = words[b + i];
            if (v >= (uint32_t)m->V) continue;
            int k = search_cdf(ca.data(), K, token_uniform(infer_seed, i, 0, TAG_INFINIT));
            z[i] = (uint16_t)k;
            ndk[k]++; nk_eff[k]++; cd[run_start[i] * K + k]++;
        }
        std::vector<int32_t> col(K);
        const int S = seg_count(n, seg_tokens, seg_max);
        std::vector<int32_t> snap(K), nd_s(K), nk_s(K), merged(K);
        // steps per iteration: one, or enough for the longest segment
        uint64_t longest = 0;
        for (int sg = 0; sg < S; ++sg)
            longest = std::max(longest, seg_start(words + b, n, sg + 1, S) - seg_start(words + b, n, sg, S));
        const uint64_t step = (S > 1 && sync_tokens) ? sync_tokens : std::max<uint64_t>(longest, 1);
        const uint64_t nsteps = std::max<uint64_t>((longest + step - 1) / step, 1);
        for (int it = 0; it < iters; ++it) for (uint64_t j = 0; j < nsteps; ++j) {
            snap = ndk;
            merged = ndk;
            for (int sg = 0; sg < S; ++sg) {
                const uint64_t a0 = seg_start(words + b, n, sg, S), a1 = seg_start(words + b, n, sg + 1, S);
                const uint64_t s0 = std::min(a1, a0 + j * step), s1 = std::min(a1, a0 + (j + 1) * step);
                nd_s = snap;
                for (int k = 0; k < K; ++k) nk_s[k] = m->nk[k] + snap[k];
                for (uint64_t i = s0; i < s1; ++i) {
                    uint32_t v = words[b + i];
                    if (v >= (uint32_t)m->V) continue;
                    int zo = z[i];
                    int32_t* c = &cd[run_start[i] * K];
                    nd_s[zo]--; nk_s[zo]--; c[zo]--;
                    for (int k = 0; k < K; ++k) col[k] = m->nkw[(size_t)v * K + k] + c[k];
                    float u = token_uniform(infer_seed, i, (uint32_t)(it + 1), TAG_INFER);
                    int zn = sample_one(K, m->scan_order, m->kpl, nd_s.data(), m->alpha.data(), col.data(), m->eta,
                                        nk_s.data(), m->veta, u, buf.data());
                    z[i] = (uint16_t)zn;
                    nd_s[zn]++; nk_s[zn]++; c[zn]++;
                }
                for (int k = 0; k < K; ++k) merged[k] += nd_s[k] - snap[k];
            }
            ndk = merged;
        }
        int64_t L = 0;
        for (uint64_t i = 0; i < n; ++i) L += words[b + i] < (uint32_t)m->V;
        double eta = (double)m->eta;
        double veta = mode ? (double)m->veta : (double)m->V * eta;
        double asum = mode ? (double)asum_f : asum_d;
        double l = 0.0;
        // getLLRest(private) - getLLRest(global): n_k terms and touched n_kw terms
        for (int k = 0; k < K; ++k) {
            if (!ndk[k]) continue;
            double x0 = mode ? (double)((float)m->nk[k] + m->veta) : (double)m->nk[k] + veta;
            double x1 = mode ? (double)((float)(m->nk[k] + ndk[k]) + m->veta) : (double)(m->nk[k] + ndk[k]) + veta;
            l += lg(x0) - lg(x1);
        }
        for (uint64_t i = 0; i < n; ++i) {
            if (run_start[i] != i) continue;
            uint32_t v = words[b + i];
            if (v >= (uint32_t)m->V) continue;
            for (int k = 0; k < K; ++k) {
                int32_t c = cd[i * K + k];
                if (!c) continue;
                int32_t g0 = m->nkw[(size_t)v * K + k];
                double y0 = mode ? (double)((float)g0 + m->eta) : (double)g0 + eta;
                double y1 = mode ? (double)((float)(g0 + c) + m->eta) : (double)(g0 + c) + eta;
                // tomoto skips zero cells; lgamma(0+eta)-lgamma(eta) is 0 either way
                l += (lg(y1) - lg(eta)) - (g0 ? (lg(y0) - lg(eta)) : 0.0);
            }
        }
        // getLLDocs(doc)
        double Ld = mode ? (double)((float)L + asum_f) : (double)L + asum;
        l -= lg(Ld) - lg(asum);
        for (int k = 0; k < K; ++k) {
            double x = mode ? (double)((float)ndk[k] + m->alpha[k]) : (double)ndk[k] + (double)m->alpha[k];
            l += lg(x) - lg((double)m->alpha[k]);
        }
        if (ll) ll[d] = l;
        if (theta) {
            for (int k = 0; k < K; ++k) theta[(size_t)d * K + k] = ((float)ndk[k] + m->alpha[k]) / ((float)L + asum_f);
        }
        if (z_out) std::copy(z.begin(), z.end(), z_out + b);
    }
}

// Brute-force fold-in LL for one finished document: copies the whole K x V
// state, adds the document, and evaluates the full getLLRest difference
// (LDAModel.hpp:931, 945-946).  Used to pin the delta form above.
double orc_infer_ll_bruteforce(void* h, const uint32_t* words, uint64_t n, const uint16_t* z) {
    Model* m = (Model*)h;
    Model tmp = *m;
    int64_t L = 0;
    std::vector<int32_t> ndk(m->K, 0);
    for (uint64_t i = 0; i < n; ++i) {
        if (words[i] >= (uint32_t)m->V) continue;
        int k = z[i];
        tmp.nkw[(size_t)words[i] * m->K + k]++;
        tmp.nk[k]++;
        ndk[k]++;
        ++L;
    }
    auto rest = [&](const Model& s) {
        double eta = (double)s.eta, veta = (double)s.V * eta;
        double r = std::lgamma(veta) * s.K;
        for (int k = 0; k < s.K; ++k) {
            r -= std::lgamma((double)s.nk[k] + veta);
            for (int v = 0; v < s.V; ++v) {
                int32_t c = s.nkw[(size_t)v * s.K + k];
                if (c) r += std::lgamma((double)c + eta) - std::lgamma(eta);
            }
        }
        return r;
    };
    double l = rest(tmp) - rest(*m);
    double asum = 0.0;
    for (int k = 0; k < m->K; ++k) asum += (double)m->alpha[k];
    l -= std::lgamma((double)L + asum) - std::lgamma(asum);
    for (int k = 0; k < m->K; ++k) l += std::lgamma((double)ndk[k] + (double)m->alpha[k]) - std::lgamma((double)m->alpha[k]);
    return l;
}

// together=True for a term-weighted model (LDAModel.hpp:852-892): the copy
// starts from the model's float n_kw / n_k and the documents' tokens are
// added onto it in document and token order (initializeDocState on tmpState);
// n_dk and the weight sums start at 0.  Then `iters` sweeps of the weighted
// schedule (with its merge: clamp and row sums) with alpha fixed, and
// ll = getLLRest(copy) - getLLRest(model) + getLLDocs(docs);
// theta = (n_dk + alpha) / (weight sum + sum alpha).
void orc_infer_together_w(void* h, const uint32_t* words, const uint64_t* off, int64_t D, const float* wts, int iters,
                          uint64_t infer_seed, int workers, int scheme, int threads, float* theta, double* ll,
                          uint16_t* z_out) {
    const Model* m = (const Model*)h;
    Model t;
    t.K = m->K; t.V = m->V; t.kpl = m->kpl; t.scan_order = m->scan_order;
    t.eta = m->eta; t.veta = m->veta; t.seed = infer_seed;
    t.alpha = m->alpha;
    t.nkw.assign((size_t)m->V * m->K, 0);
    t.nk.assign(m->K, 0);
    t.chunk_cf.assign(m->V, 0);
    for (uint32_t v : m->words) if (v < (uint32_t)m->V) ++t.chunk_cf[v];
    t.optim_interval = 0;
    t.eta_kv = m->eta_kv; t.etasum = m->etasum; t.etasum_d = m->etasum_d;
    orc_set_corpus_w(&t, words, off, D, nullptr, wts);
    t.nkw_f = m->nkw_f;
    t.nk_f = m->nk_f;
    std::fill(t.ndk_f.begin(), t.ndk_f.end(), 0.0f);
    for (int64_t d = 0; d < D; ++d) {
        float sw = 0.0f;
        for (uint64_t g = off[d]; g < off[d + 1]; ++g) {
            const uint32_t v = words[g];
            if (v >= (uint32_t)t.V) continue;
            add_token_w(t, d, v, t.z[g], t.wt[g], false, t.nk_f.data());
            sw += t.wt[g];
        }
        t.docw[d] = sw;
    }
    orc_train_w(&t, iters, workers, threads, scheme);
    double docs = 0, rest = 0, docs0 = 0, rest0 = 0;
    loglik_w(t, &docs, &rest);
    loglik_w(*m, &docs0, &rest0);
    *ll = rest - rest0 + docs;
    if (z_out) std::copy(t.z.begin(), t.z.end(), z_out);
    float asum = 0.0f;
    for (int k = 0; k < m->K; ++k) asum += m->alpha[k];
    for (int64_t d = 0; d < D; ++d)
        for (int k = 0; k < m->K; ++k)
            theta[(size_t)d * m->K + k] = (t.ndk_f[(size_t)d * m->K + k] + m->alpha[k]) / (t.docw[d] + asum);
}

// Fold-in (together=False) of term-weighted models and of models with word
// priors (LDAModel.hpp:920-953 with initializeDocState 776-813 and
// updateStateWithDoc 764-775).  Per document, on a private copy of the
// model's n_kw / n_k:
//   init  z_i ~ Discrete(alpha), or with priors z_i ~ sampleFromDiscrete of
//         the word's eta column (u keyed by infer_seed, the token position i
//         inside the document, 0, TAG_INFINIT); then addWordTo<+1> with the
//         token's weight (1 for ONE) in token order;
//   iters sweeps over the tokens in order: remove (weighted: max(x - w, 0)),
//         getZLikelihoods with eta_kv / sum_v eta_kv, sample (u keyed by i,
//         it + 1, TAG_INFER), add;
//   ll = getLLRest(private) - getLLRest(global) + getLLDocs(doc) in FP64;
//   theta = (n_dk + alpha) / (sumWordWeight + sum alpha), the weight sum
//         accumulated in float in token order (updateSumWordWeight, LDA.h:113-116).
// wts: per-token weights (IDF / PMI models, whose float counts are set), or
// NULL (int counts).  faithful = 1 keeps tomoto's full K x V private copy and
// evaluates getLLRest over all of it; faithful = 0 keeps only the document's
// word columns (one per run of equal word ids) and sums the LL over the cells
// and topics the document changed (an unchanged cell contributes exactly 0).
// faithful = 2 is faithful = 1 with tomoto's own evaluation of the LL
// (LDAModel.hpp:652-664, 925-946): the table lgamma, getLLRest(global)
// computed once per call, and the private copy scanned topic by topic over
// every word, skipping empty cells.  It exists to time tomoto's form on the
// CPU; its LL carries the table's interpolation error.
void orc_infer_gen_seg(void* h, const uint32_t* words, const uint64_t* off, int64_t D, int iters, uint64_t infer_seed,
                       const float* wts, int faithful, uint32_t seg_tokens, int seg_max, float* theta, double* ll,
                       uint16_t* z_out);

void orc_infer_gen(void* h, const uint32_t* words, const uint64_t* off, int64_t D, int iters, uint64_t infer_seed,
                   const float* wts, int faithful, float* theta, double* ll, uint16_t* z_out) {
    orc_infer_gen_seg(h, words, off, D, iters, infer_seed, wts, faithful, 0, 1, theta, ll, z_out);
}

// Segmented rule for these models (the device's, for long documents; the
// integer fold-in's rule of orc_infer_seg restated on running counts): a
// document of S > 1 segments starts every iteration with each segment on a
// copy of n_dk and n_k as they were when the iteration started (the
// columns are the segments' own: a word's run lies in one segment); after the
// iteration n_dk = snap + sum_s (n_dk,s - snap) and n_k likewise, the
// differences added in segment order (in float for term-weighted counts).
// S == 1 samples the document sequentially as above.  The initial topics are
// added in token order before the first iteration, as for S == 1.
void orc_infer_gen_seg(void* h, const uint32_t* words, const uint64_t* off, int64_t D, int iters, uint64_t infer_seed,
                       const float* wts, int faithful, uint32_t seg_tokens, int seg_max, float* theta, double* ll,
                       uint16_t* z_out) {
    Model* m = (Model*)h;
    const int K = m->K, V = m->V;
    const bool W = wts != nullptr;
    const bool P = !m->eta_kv.empty();
    std::vector<float> ca(K);
    float asf = 0.0f;
    for (int k = 0; k < K; ++k) { asf = asf + m->alpha[k]; ca[k] = asf; }
    double asd = 0.0;
    for (int k = 0; k < K; ++k) asd += (double)m->alpha[k];
    std::vector<float> buf(64 * device_kpl(K) + K + 64);
    const double eta = (double)m->eta, veta = (double)V * eta;
    auto es_of = [&](int k) { return P ? m->etasum_d[k] : veta; };
    auto e_of = [&](uint32_t v, int k) { return P ? (double)m->eta_kv[(size_t)v * K + k] : eta; };
    // frozen counts as double (int or float model)
    auto g_nk = [&](int k) { return W ? (double)m->nk_f[k] : (double)m->nk[k]; };
    auto g_kw = [&](uint32_t v, int k) { return W ? (double)m->nkw_f[(size_t)v * K + k] : (double)m->nkw[(size_t)v * K + k]; };
    // getLLRest of a K x V state with the table lgamma, topic by topic over every word
    auto rest_lut = [&](auto&& nk_of, auto&& kw_of) {
        auto lgt = [](double x) { return (double)lut_lgamma().get((float)x); };
        double r = 0.0;
        for (int k = 0; k < K; ++k) {
            r += lgt(es_of(k)) - lgt(nk_of(k) + es_of(k));
            for (uint32_t v = 0; v < (uint32_t)V; ++v) {
                const double c = kw_of(v, k);
                if (c != 0.0) r += lgt(c + e_of(v, k)) - lgt(e_of(v, k));
            }
        }
        return r;
    };
    const double grest = faithful == 2 ? rest_lut(g_nk, g_kw) : 0.0;
    for (int64_t d = 0; d < D; ++d) {
        const uint64_t b = off[d], n = off[d + 1] - b;
        std::vector<uint16_t> z(n, 0xFFFF);
        std::vector<uint64_t> run(n);  // index of the first token of each token's run
        for (uint64_t i = 0; i < n; ++i) run[i] = (i && words[b + i] == words[b + i - 1]) ? run[i - 1] : i;
        // private state: n_dk, n_k, and the columns (run-start x K, or the full V x K)
        std::vector<float> ndf(K, 0.0f), nkf;
        std::vector<int32_t> ndi(K, 0), nki;
        std::vector<float> cf;
        std::vector<int32_t> ci;
        if (W) nkf = m->nk_f; else nki = m->nk;
        if (faithful) { if (W) cf = m->nkw_f; else ci = m->nkw; }
        else if (W) cf.assign((size_t)n * K, 0.0f);
        else ci.assign((size_t)n * K, 0);
        auto cell_f = [&](uint64_t i, uint32_t v) { return faithful ? &cf[(size_t)v * K] : &cf[run[i] * K]; };
        auto cell_i = [&](uint64_t i, uint32_t v) { return faithful ? &ci[(size_t)v * K] : &ci[run[i] * K]; };
        if (!faithful) {
            for (uint64_t i = 0; i < n; ++i) {
                const uint32_t v = words[b + i];
                if (run[i] != i || v >= (uint32_t)V) continue;
                for (int k = 0; k < K; ++k) {
                    if (W) cf[i * K + k] = m->nkw_f[(size_t)v * K + k];
                    else ci[i * K + k] = m->nkw[(size_t)v * K + k];
                }
            }
        }
        float sw = 0.0f;
        int64_t L = 0;
        for (uint64_t i = 0; i < n; ++i) {
            const uint32_t v = words[b + i];
            if (v >= (uint32_t)V) continue;
            const float u = token_uniform(infer_seed, i, 0, TAG_INFINIT);
            const int k = P ? sample_from_discrete(m->ecol(v), K, u) : search_cdf(ca.data(), K, u);
            z[i] = (uint16_t)k;
            if (W) {
                const float w = wts[b + i];
                ndf[k] += w; nkf[k] += w; cell_f(i, v)[k] += w;
                sw += w;
            } else {
                ndi[k]++; nki[k]++; cell_i(i, v)[k]++;
            }
            ++L;
        }
        // one segment's tokens [s0, s1) on the state (nd, nk)
        auto sample_seg = [&](int it, uint64_t s0, uint64_t s1, float* ndF, float* nkF, int32_t* ndI, int32_t* nkI) {
            for (uint64_t i = s0; i < s1; ++i) {
                const uint32_t v = words[b + i];
                if (v >= (uint32_t)V) continue;
                const int zo = z[i];
                const float u = token_uniform(infer_seed, i, (uint32_t)(it + 1), TAG_INFER);
                int zn;
                if (W) {
                    const float w = wts[b + i];
                    float* c = cell_f(i, v);
                    ndF[zo] = sub_clamp(ndF[zo], w); nkF[zo] = sub_clamp(nkF[zo], w); c[zo] = sub_clamp(c[zo], w);
                    zn = sample_one_w(*m, ndF, c, nkF, u, buf.data(), m->ecol(v));
                    ndF[zn] += w; nkF[zn] += w; c[zn] += w;
                } else {
                    int32_t* c = cell_i(i, v);
                    ndI[zo]--; nkI[zo]--; c[zo]--;
                    zn = sample_one(K, m->scan_order, m->kpl, ndI, m->alpha.data(), c, m->eta, nkI, m->veta, u,
                                    buf.data(), m->ecol(v), m->esum());
                    ndI[zn]++; nkI[zn]++; c[zn]++;
                }
                z[i] = (uint16_t)zn;
            }
        };
        const int S = faithful ? 1 : seg_count(n, seg_tokens, seg_max);
        std::vector<float> sndf, snkf, accf, accg, segf, segg;
        std::vector<int32_t> sndi, snki, acci, accj, segi, segj;
        for (int it = 0; it < iters; ++it) {
            if (S == 1) {
                sample_seg(it, 0, n, ndf.data(), nkf.data(), ndi.data(), nki.data());
                continue;
            }
            sndf = ndf; snkf = nkf; sndi = ndi; snki = nki;
            accf = ndf; accg = nkf; acci = ndi; accj = nki;
            for (int sg = 0; sg < S; ++sg) {
                const uint64_t s0 = seg_start(words + b, n, sg, S), s1 = seg_start(words + b, n, sg + 1, S);
                segf = sndf; segg = snkf; segi = sndi; segj = snki;
                sample_seg(it, s0, s1, segf.data(), segg.data(), segi.data(), segj.data());
                for (int k = 0; k < K; ++k) {
                    if (W) {
                        accf[k] = accf[k] + (segf[k] - sndf[k]);
                        accg[k] = accg[k] + (segg[k] - snkf[k]);
                    } else {
                        acci[k] += segi[k] - sndi[k];
                        accj[k] += segj[k] - snki[k];
                    }
                }
            }
            ndf = accf; nkf = accg; ndi = acci; nki = accj;
        }
        // getLLRest(private) - getLLRest(global)
        double l = 0.0;
        if (faithful == 2) {
            l = rest_lut([&](int k) { return W ? (double)nkf[k] : (double)nki[k]; },
                         [&](uint32_t v, int k) {
                             return W ? (double)cf[(size_t)v * K + k] : (double)ci[(size_t)v * K + k];
                         }) - grest;
        } else {
            for (int k = 0; k < K; ++k) {
                const double x0 = g_nk(k), x1 = W ? (double)nkf[k] : (double)nki[k];
                if (faithful || x0 != x1) l += std::lgamma(x0 + es_of(k)) - std::lgamma(x1 + es_of(k));
            }
            auto cell_term = [&](double c, double e) { return c != 0.0 ? std::lgamma(c + e) - std::lgamma(e) : 0.0; };
            if (faithful) {
                for (uint32_t v = 0; v < (uint32_t)V; ++v)
                    for (int k = 0; k < K; ++k) {
                        const double c1 = W ? (double)cf[(size_t)v * K + k] : (double)ci[(size_t)v * K + k];
                        l += cell_term(c1, e_of(v, k)) - cell_term(g_kw(v, k), e_of(v, k));
                    }
            } else {
                for (uint64_t i = 0; i < n; ++i) {
                    const uint32_t v = words[b + i];
                    if (run[i] != i || v >= (uint32_t)V) continue;
                    for (int k = 0; k < K; ++k) {
                        const double c0 = g_kw(v, k), c1 = W ? (double)cf[i * K + k] : (double)ci[i * K + k];
                        if (c0 != c1) l += cell_term(c1, e_of(v, k)) - cell_term(c0, e_of(v, k));
                    }
                }
            }
        }
        // getLLDocs(doc)
        const double Ld = W ? (double)sw : (double)L;
        l -= std::lgamma(Ld + asd) - std::lgamma(asd);
        for (int k = 0; k < K; ++k) {
            const double x = W ? (double)ndf[k] : (double)ndi[k];
            if (x != 0.0) l += std::lgamma(x + (double)m->alpha[k]) - std::lgamma((double)m->alpha[k]);
        }
        if (ll) ll[d] = l;
        if (theta) {
            const float den = (W ? sw : (float)L) + asf;
            for (int k = 0; k < K; ++k)
                theta[(size_t)d * K + k] = ((W ? ndf[k] : (float)ndi[k]) + m->alpha[k]) / den;
        }
        if (z_out) std::copy(z.begin(), z.end(), z_out + b);
    }
}

}  // extern "C"
