"""Pins of the CPU oracle (no GPU needed).

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong association, an off-by-one in the search, a wrong lgamma tail)
fails it:
  * Philox4x32-10 against the published Random123 known-answer vectors;
  * weights against the reference expression evaluated in numpy float32;
  * tomoto-order scan, search, shuffle order and LUT lgamma / digamma against
    the reference's own headers compiled in place (oracle/_ref, when built);
  * device-order scan against a float64 cumsum within rounding bounds;
  * FP64 log-likelihood against scipy.special.gammaln on the closed form;
  * the Gibbs chain against the exact collapsed posterior (enumeration);
  * the alpha fixed point against a known alpha;
  * fold-in's delta log-likelihood against the full K x V copy.
"""
import ctypes
import itertools
import os

import numpy as np
import pytest
from scipy.special import gammaln

from oracle import oracle as O
from tomotopy_amd import corpus, synth

REF_LIB = os.path.join(os.path.dirname(O.__file__), "_ref", "libtomoto_ref.so")


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(REF_LIB):
        pytest.skip("reference utilities (oracle/_ref) not built here")
    r = ctypes.CDLL(REF_LIB)
    r.ref_prefix_sum.argtypes = [ctypes.c_void_p, ctypes.c_int]
    r.ref_sample_acc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_float]
    r.ref_sample_acc.restype = ctypes.c_int
    for n in ("ref_lgammaT", "ref_digammaT"):
        getattr(r, n).argtypes = [ctypes.c_float]
        getattr(r, n).restype = ctypes.c_float
    r.ref_for_shuffled.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p]
    r.ref_vocab_order.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                  ctypes.c_uint64, ctypes.c_void_p]
    return r


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ------------------------------------------------------------------ Philox --
def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds
    assert list(O.philox([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert list(O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0])) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_token_uniform_is_24_bit_and_in_range():
    us = np.array([O.token_uniform(7, g, 3) for g in range(2000)], dtype=np.float32)
    assert us.min() >= 0.0 and us.max() < 1.0
    assert np.all(us * 2**24 == np.floor(us * 2**24))
    assert 0.45 < us.mean() < 0.55


# ----------------------------------------------------------------- weights --
@pytest.mark.parametrize("K", [1, 3, 20, 1000])
def test_weights_closed_form(K):
    rng = np.random.default_rng(K)
    ndk = rng.integers(0, 500, K).astype(np.int32)
    nkw = rng.integers(0, 100000, K).astype(np.int32)
    nk = rng.integers(0, 10**7, K).astype(np.int32)
    alpha = (rng.random(K) + 0.01).astype(np.float32)
    eta = np.float32(0.01)
    veta = np.float32(eta * np.float32(100000))
    # LDAModel.hpp:280-282: ((n_dk + alpha) * (n_kw + eta)) / (n_k + V*eta), float32
    exp = ((ndk.astype(np.float32) + alpha) * (nkw.astype(np.float32) + eta)) / (nk.astype(np.float32) + veta)
    got = O.weights(ndk, alpha, nkw, eta, nk, veta)
    assert got.tobytes() == exp.astype(np.float32).tobytes()
    wrong = (ndk.astype(np.float32) + alpha) * ((nkw.astype(np.float32) + eta) / (nk.astype(np.float32) + veta))
    if K >= 20:
        assert got.tobytes() != wrong.astype(np.float32).tobytes()


def test_markstein_quotient_matches_ieee_division():
    # the device computes a/b as q0=a*y, r=fma(-q0,b,a), q=fma(r,y,q0) with y=RN(1/b)
    assert O.markstein_mismatches(3_000_000, 11, 1e-4, 1e13, 1.0, 1e9) == 0


# ------------------------------------------------------------------- scans --
@pytest.mark.parametrize("n", list(range(1, 18)) + [20, 63, 64, 100, 1000, 1001])
def test_prefix_sum_tomoto_matches_reference(ref, n):
    rng = np.random.default_rng(n)
    for trial in range(20):
        x = (rng.random(n) * 10.0 ** rng.integers(-6, 4, n)).astype(np.float32)
        a = x.copy()
        ref.ref_prefix_sum(_p(a), n)
        b = O.prefix_sum_tomoto(x)
        assert a.tobytes() == b.tobytes()


@pytest.mark.parametrize("kpl", [1, 2, 4, 8, 16, 32, 64, 128, 512])
def test_prefix_sum_device_is_a_prefix_sum(kpl):
    rng = np.random.default_rng(kpl)
    x = (rng.random(64 * kpl) * 10.0 ** rng.integers(-3, 3, 64 * kpl)).astype(np.float32)
    c = O.prefix_sum_device(x, kpl)
    exact = np.cumsum(x.astype(np.float64))
    # every partial sum of n non-negative terms is within n*2^-24 relative
    assert np.all(np.abs(c - exact) <= exact * (64 * kpl) * 2.0**-24 + 1e-30)
    assert np.all(np.diff(c[:kpl]) >= 0)
    # single non-zero entries are carried exactly to every later position
    for pos in (0, 17, 64 * kpl - 1):
        e = np.zeros(64 * kpl, np.float32)
        e[pos] = 3.5
        ce = O.prefix_sum_device(e, kpl)
        assert np.all(ce[:pos] == 0) and np.all(ce[pos:] == 3.5)


# ------------------------------------------------------------------ search --
def test_search_matches_reference_and_brute_force(ref):
    rng = np.random.default_rng(1)
    for K in (1, 2, 7, 8, 31, 32, 33, 100, 1000):
        for trial in range(200):
            w = rng.random(K).astype(np.float32)
            w[rng.random(K) < 0.3] = 0.0
            if w.sum() == 0:
                w[0] = 1.0
            cdf = O.prefix_sum_tomoto(w)
            u = np.float32(rng.random()) if trial % 5 else np.float32(1.0 - 2**-24)
            got = O.search(cdf, K, u)
            r = np.float32(u * cdf[K - 1])
            hits = np.nonzero(r < cdf)[0]
            brute = int(hits[0]) if len(hits) else K - 1
            assert got == brute
            assert got == ref.ref_sample_acc(_p(cdf), K, float(u))
    # r equal to a cdf value moves past it (strict <)
    cdf = np.array([1.0, 2.0, 2.0, 4.0], np.float32)
    assert O.search(cdf, 4, 0.5) == 3
    assert O.search(cdf, 4, 0.25) == 1


def test_shuffled_order_matches_reference(ref):
    for N in (1, 2, 5, 100, 65537, 131074, 200003):
        for seed in (0, 1, 17, 0xDEADBEEF):
            o = O.shuffled_order(N, seed)
            a = np.zeros(N, np.uint32)
            ref.ref_for_shuffled(N, seed, _p(a))
            assert np.array_equal(o, a)
            assert len(np.unique(o)) == N


# --------------------------------------------------------------------- LUT --
def test_lgamma_lut_matches_reference_bit_exact(ref):
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.random(5000) * 1200, rng.random(500) * 0.002, rng.random(500) * 5000,
                         [0.0001, 0.001, 0.5, 1, 100, 101, 1100.875, 1200, 0.01, 0.1]]).astype(np.float32)
    for x in xs:
        assert np.float32(O.lgamma_lut(x)).tobytes() == np.float32(ref.ref_lgammaT(x)).tobytes(), x


def test_lut_within_interpolation_bound():
    # linear interpolation error <= h^2/8 * max|f''| on a step of width h
    for x in np.linspace(0.002, 1100, 3001):
        h = 1 / 1024 if x < 1 else (1 / 64 if x < 101 else 1 / 8)
        bound = h * h / 8 * (1 / (x * x) + 1 / x + 1e-3) + 3e-7 * max(1.0, abs(gammaln(x)))
        assert abs(O.lgamma_lut(x) - gammaln(x)) <= bound + 1e-6, x


def test_digamma_lut_close_to_reference(ref):
    for x in np.concatenate([np.linspace(0.0005, 1200, 4000), [0.01, 0.05, 0.1, 1.0, 2.0]]).astype(np.float32):
        a, b = O.digamma_lut(x), ref.ref_digammaT(x)
        assert abs(a - b) <= 2e-6 * max(1.0, abs(b)), x


# ------------------------------------------------------------- vocabulary --
def test_vocab_order_matches_reference(ref):
    rng = np.random.default_rng(3)
    for trial in range(30):
        n = int(rng.integers(1, 300))
        cf = rng.integers(0, 20, n).astype(np.uint64)
        df = np.minimum(cf, rng.integers(0, 10, n)).astype(np.uint64)
        min_cf, min_df, rm_top = int(rng.integers(0, 4)), int(rng.integers(0, 3)), int(rng.integers(0, 5))
        order, real_v = corpus.vocab_order(cf, df, min_cf, min_df, rm_top)
        r = np.zeros(n, np.uint32)
        ref.ref_vocab_order(n, _p(cf), _p(df), min_cf, min_df, rm_top, _p(r))
        assert np.array_equal(order, r)


# ---------------------------------------------------------------- full chain --
def _collapsed_log_posterior(z, docs, K, V, alpha, eta):
    """log P(z | w) up to a constant, Dirichlet-multinomial closed form (gammaln)."""
    lp = 0.0
    nkv = np.zeros((K, V))
    for d, ws in enumerate(docs):
        ndk = np.bincount(z[d], minlength=K)
        lp += gammaln(K * alpha) - gammaln(len(ws) + K * alpha) + np.sum(gammaln(ndk + alpha) - gammaln(alpha))
        for w, k in zip(ws, z[d]):
            nkv[k, w] += 1
    lp += np.sum(gammaln(V * eta) - gammaln(nkv.sum(1) + V * eta))
    lp += np.sum(gammaln(nkv + eta) - gammaln(eta))
    return lp


@pytest.mark.slow
@pytest.mark.parametrize("workers", [1, 2])
def test_chain_marginals_match_exact_posterior(workers):
    """Tiny corpus, K=2: enumerate all 2^N states; the chain's empirical
    marginals of z must match the exact posterior marginals."""
    docs = [[0, 0, 1], [1, 2, 2], [0, 2]]
    K, V, alpha, eta = 2, 3, 0.5, 0.3
    N = sum(len(d) for d in docs)
    lens = [len(d) for d in docs]
    post = np.zeros((N, K))
    tot = 0.0
    for bits in itertools.product(range(K), repeat=N):
        z = np.array(bits)
        zs, pos = [], 0
        for L in lens:
            zs.append(z[pos:pos + L]); pos += L
        p = np.exp(_collapsed_log_posterior(zs, docs, K, V, alpha, eta))
        tot += p
        post[np.arange(N), z] += p
    post /= tot
    words = np.array([w for d in docs for w in d], np.uint32)  # already sorted per doc
    off = np.array([0, 3, 6, 8], np.uint64)
    m = O.Model(K, V, alpha, eta, seed=123, scan_order=1)
    m.set_optim(0, 0)
    m.set_corpus(words, off)
    S = 30000
    counts = np.zeros((N, K))
    for s in range(S):
        m.train(1, workers=workers)
        z = m.state()[0]
        counts[np.arange(N), z] += 1
    emp = counts / S
    assert np.max(np.abs(emp - post)) < 0.02, (emp, post)


# ------------------------------------------------------------ count checks --
def test_oracle_counts_stay_consistent(small_corpus):
    ws, offs, V = small_corpus
    m = O.Model(20, V, 0.1, 0.01, seed=1, scan_order=1)
    m.set_corpus(ws, offs)
    m.train(12, workers=4)
    z, ndk, nkw, nk = m.state()
    D = len(offs) - 1
    doc_of = np.repeat(np.arange(D), np.diff(offs).astype(np.int64))
    assert np.array_equal(ndk.sum(1), np.diff(offs).astype(np.int64))
    assert np.array_equal(nkw.sum(0), nk)
    rec = np.zeros_like(ndk)
    np.add.at(rec, (doc_of, z.astype(np.int64)), 1)
    assert np.array_equal(rec, ndk)
    rec2 = np.zeros_like(nkw)
    np.add.at(rec2, (ws.astype(np.int64), z.astype(np.int64)), 1)
    assert np.array_equal(rec2, nkw)


# -------------------------------------------------------------------- LL ----
def test_loglik_fp64_matches_scipy(small_corpus):
    ws, offs, V = small_corpus
    K = 20
    rng = np.random.default_rng(5)
    alpha = (rng.random(K) * 0.3 + 0.02).astype(np.float32)
    z = rng.integers(0, K, len(ws)).astype(np.uint16)
    m = O.Model(K, V, alpha, 0.01, seed=1)
    m.set_corpus(ws, offs, z)
    ld, lr = m.loglik(0)
    D = len(offs) - 1
    doc_of = np.repeat(np.arange(D), np.diff(offs).astype(np.int64))
    ndk = np.zeros((D, K)); np.add.at(ndk, (doc_of, z.astype(np.int64)), 1)
    nkw = np.zeros((V, K)); np.add.at(nkw, (ws.astype(np.int64), z.astype(np.int64)), 1)
    a = alpha.astype(np.float64)
    eta = float(np.float32(0.01))
    L = np.diff(offs).astype(np.float64)
    exp_docs = np.sum(gammaln(a.sum()) - gammaln(L + a.sum())) + np.sum(gammaln(ndk + a) - gammaln(a))
    exp_rest = K * gammaln(V * eta) - np.sum(gammaln(nkw.sum(0) + V * eta)) + np.sum(gammaln(nkw + eta) - gammaln(eta))
    assert abs(ld - exp_docs) <= 1e-9 * abs(exp_docs)
    assert abs(lr - exp_rest) <= 1e-9 * abs(exp_rest)
    # tomoto's LUT mode stays close to FP64
    td, tr = m.loglik(1)
    assert abs((td + tr) - (ld + lr)) <= 2e-4 * abs(ld + lr)


# ------------------------------------------------------------------ alpha ---
def test_alpha_fixed_point_recovers_known_alpha():
    rng = np.random.default_rng(0)
    K, D, V = 8, 4000, 50
    true = np.array([0.05, 0.1, 0.2, 0.3, 0.5, 0.8, 1.0, 1.5])
    lens = rng.integers(50, 150, D)
    words, zs = [], []
    for d in range(D):
        th = rng.dirichlet(true)
        z = rng.choice(K, lens[d], p=th)
        zs.append(z)
        words.append(np.sort(rng.integers(0, V, lens[d])))
    off = np.zeros(D + 1, np.uint64)
    off[1:] = np.cumsum(lens)
    m = O.Model(K, V, 0.3, 0.01, seed=1)
    m.set_corpus(np.concatenate(words).astype(np.uint32), off, np.concatenate(zs).astype(np.uint16))
    for _ in range(20):
        m.optimize_alpha()
    np.testing.assert_allclose(m.alpha(), true, rtol=0.12)


# ------------------------------------------------------------------ infer ---
def test_infer_delta_ll_matches_full_copy(small_corpus):
    ws, offs, V = small_corpus
    K = 20
    m = O.Model(K, V, 0.1, 0.01, seed=2)
    m.set_corpus(ws, offs)
    m.train(5, workers=1)
    uw, uoff = synth.unseen_docs(5, 6, V, 20, np.log(60.0), 0.7, 3, 300, zipf_s=None)
    uw, uoff = corpus.csr_from_arrays(uw, uoff, np.arange(V))
    theta, ll, z = m.infer(uw, uoff, 10, infer_seed=0, mode=0)
    for d in range(len(uoff) - 1):
        b, e = int(uoff[d]), int(uoff[d + 1])
        bf = m.infer_ll_bruteforce(uw[b:e], z[b:e])
        assert abs(ll[d] - bf) <= 1e-8 * max(1.0, abs(bf))
        L = e - b
        nd = np.bincount(z[b:e], minlength=K)
        np.testing.assert_allclose(theta[d], (nd + np.float32(0.1)) / (L + np.float32(0.1) * K), rtol=1e-6)


def test_infer_document_alone_equals_in_batch(small_corpus):
    ws, offs, V = small_corpus
    m = O.Model(20, V, 0.1, 0.01, seed=2)
    m.set_corpus(ws, offs)
    m.train(3, workers=1)
    uw, uoff = synth.unseen_docs(8, 5, V, 20, np.log(60.0), 0.7, 3, 300, zipf_s=None)
    uw, uoff = corpus.csr_from_arrays(uw, uoff, np.arange(V))
    th, ll, z = m.infer(uw, uoff, 8)
    b, e = int(uoff[2]), int(uoff[3])
    th1, ll1, z1 = m.infer(uw[b:e], np.array([0, e - b], np.uint64), 8)
    assert np.array_equal(th1[0], th[2]) and ll1[0] == ll[2] and np.array_equal(z1, z[b:e])


# -------------------------------------------------------- segmented infer ---
def test_segment_rule():
    """S is a power of two <= seg_max, segments tile the document, every
    segment starts at a word-run start, and no word straddles two segments."""
    rng = np.random.default_rng(3)
    for trial in range(300):
        n = int(rng.integers(0, 400))
        w = np.sort(rng.integers(0, max(1, int(rng.integers(1, 60))), n)).astype(np.uint32)
        T = int(rng.choice([0, 1, 3, 16, 50, 200]))
        M = int(rng.choice([1, 2, 4, 8]))
        st = O.seg_bounds(w, T, M)
        S = len(st) - 1
        need = 1 if (T == 0 or n <= T) else -(-n // T)
        expect = 1
        while expect < need and expect < M:
            expect *= 2
        assert S == expect
        assert st[0] == 0 and st[-1] == n and np.all(np.diff(st.astype(np.int64)) >= 0)
        for x in st[1:-1]:
            x = int(x)
            assert x == n or x == 0 or w[x] != w[x - 1]
        for i in range(1, S):
            x = int(st[i])
            nominal = (n * i) // S
            assert x >= nominal and (x == nominal or np.all(w[nominal:x] == w[max(nominal - 1, 0)]))


def test_segmented_infer_reduces_to_sequential(small_corpus):
    """seg_tokens == 0, or documents no longer than seg_tokens: bit-identical to
    the sequential fold-in."""
    ws, offs, V = small_corpus
    m = O.Model(20, V, 0.1, 0.01, seed=2)
    m.set_corpus(ws, offs)
    m.train(3, workers=1)
    uw, uoff = synth.unseen_docs(9, 12, V, 20, np.log(60.0), 0.8, 3, 400, zipf_s=None)
    uw, uoff = corpus.csr_from_arrays(uw, uoff, np.arange(V))
    th0, ll0, z0 = m.infer(uw, uoff, 6)
    th1, ll1, z1 = m.infer(uw, uoff, 6, seg_tokens=0, seg_max=8)
    assert np.array_equal(z0, z1) and np.array_equal(th0, th1) and np.array_equal(ll0, ll1)
    longest = int(np.diff(uoff.astype(np.int64)).max())
    th2, ll2, z2 = m.infer(uw, uoff, 6, seg_tokens=longest, seg_max=8)
    assert np.array_equal(z0, z2) and np.array_equal(ll0, ll2)


def _py_segmented_infer(nkw, nk, alpha, eta, V, doc, iters, seed, seg_tokens, seg_max, sync=0):
    """Pure-Python restatement of the segmented fold-in of one document, built
    from the separately pinned primitives (weights, device scan, search,
    Philox uniforms): per step (the whole iteration when sync == 0, else
    `sync` tokens of every segment) each segment samples against the step's
    starting n_dk plus its own changes, and the changes are summed."""
    K = len(alpha)
    kpl = O.device_kpl(K)
    veta = np.float32(np.float32(eta) * np.float32(V))
    ca = np.cumsum(alpha.astype(np.float32), dtype=np.float32)
    n = len(doc)
    z = np.full(n, 0xFFFF, np.int64)
    for i in range(n):
        if doc[i] < V:
            z[i] = O.search(ca, K, O.token_uniform(seed, i, 0, "infinit"))
    st = O.seg_bounds(doc, seg_tokens, seg_max)
    S = len(st) - 1
    longest = max(int(st[s + 1]) - int(st[s]) for s in range(S))
    step = sync if (S > 1 and sync) else max(longest, 1)
    steps = [(it, q) for it in range(1, iters + 1) for q in range(max(1, -(-longest // step)))]
    for it, q in steps:
        snap = np.bincount(z[z != 0xFFFF], minlength=K).astype(np.int64)
        delta = np.zeros(K, np.int64)
        for s in range(S):
            nd = snap.copy()
            a0, a1 = int(st[s]), int(st[s + 1])
            for i in range(min(a1, a0 + q * step), min(a1, a0 + (q + 1) * step)):
                v = int(doc[i])
                if v >= V:
                    continue
                zo = int(z[i])
                nd[zo] -= 1
                same = [j for j in range(n) if doc[j] == v and j != i]  # c_d: the document's other tokens of v
                col = nkw[v].astype(np.int64).copy()
                for j in same:
                    col[z[j]] += 1
                w = O.weights(nd, alpha, col, eta, nk + nd, veta)
                cdf = O.prefix_sum_device(w, kpl)
                zn = O.search(cdf, K, O.token_uniform(seed, i, it, "infer"))
                z[i] = zn
                nd[zn] += 1
            delta += nd - snap
    return z


def test_segmented_infer_matches_python_restatement():
    """The oracle's segmented fold-in equals a direct Python restatement of the
    rule (per-iteration snapshot, per-segment private n_dk, summed changes)."""
    K, V = 6, 30
    rng = np.random.default_rng(11)
    D = 40
    lens = rng.integers(20, 60, D)
    words = [np.sort(rng.integers(0, V, L)) for L in lens]
    off = np.zeros(D + 1, np.uint64)
    off[1:] = np.cumsum(lens)
    alpha = np.float32(0.2)
    m = O.Model(K, V, alpha, 0.05, seed=3)
    m.set_corpus(np.concatenate(words).astype(np.uint32), off)
    m.train(4, workers=1)
    _, _, nkw, nk = m.state()
    al = np.full(K, alpha, np.float32)
    for seg_tokens, seg_max, n, sync in ((10, 4, 37, 0), (20, 2, 40, 0), (0, 8, 25, 0), (3, 8, 9, 0),
                                         (10, 4, 37, 3), (20, 2, 40, 7), (3, 8, 9, 1), (10, 4, 37, 64)):
        doc = np.sort(rng.integers(0, V, n)).astype(np.uint32)
        o = np.array([0, n], np.uint64)
        _, _, z = m.infer(doc, o, 3, infer_seed=7, seg_tokens=seg_tokens, seg_max=seg_max, seg_sync=sync)
        zp = _py_segmented_infer(nkw, nk.astype(np.int64), al, 0.05, V, doc, 3, 7, seg_tokens, seg_max, sync)
        assert np.array_equal(z.astype(np.int64), zp), (seg_tokens, seg_max, sync)
    # the steps matter: merging every 3 tokens samples other topics than once per iteration
    doc = np.sort(rng.integers(0, V, 60)).astype(np.uint32)
    o = np.array([0, 60], np.uint64)
    _, _, za = m.infer(doc, o, 3, infer_seed=7, seg_tokens=10, seg_max=4, seg_sync=0)
    _, _, zb = m.infer(doc, o, 3, infer_seed=7, seg_tokens=10, seg_max=4, seg_sync=3)
    assert not np.array_equal(za, zb)
    # LL and theta of a segmented result stay the closed forms of its z
    th, ll, z = m.infer(doc, o, 5, seg_tokens=2, seg_max=4)
    nd = np.bincount(z[z != 0xFFFF], minlength=K)
    np.testing.assert_allclose(th[0], (nd + alpha) / (len(doc) + alpha * K), rtol=1e-6)
    bf = m.infer_ll_bruteforce(doc, z)
    assert abs(ll[0] - bf) <= 1e-8 * max(1.0, abs(bf))


def test_segmented_infer_same_ll_per_word_as_sequential(small_corpus):
    """The segmented sampler targets the same posterior: over many long
    documents its mean log-likelihood per word matches the sequential
    sampler's within 1%."""
    ws, offs, V = small_corpus
    m = O.Model(20, V, 0.1, 0.01, seed=5)
    m.set_corpus(ws, offs)
    m.train(30, workers=1)
    uw, uoff = synth.unseen_docs(21, 40, V, 20, np.log(600.0), 0.3, 300, 1200, zipf_s=None)
    uw, uoff = corpus.csr_from_arrays(uw, uoff, np.arange(V))
    n = int(np.count_nonzero(uw < V))
    _, ll_seq, _ = m.infer(uw, uoff, 30, infer_seed=4)
    _, ll_seg, _ = m.infer(uw, uoff, 30, infer_seed=4, seg_tokens=64, seg_max=8)
    a, b = ll_seq.sum() / n, ll_seg.sum() / n
    print("ll_per_word sequential %.5f segmented %.5f" % (a, b))
    assert abs(a - b) <= 0.01 * abs(a)


def test_infer_together_closed_forms():
    """together=True: from the returned topics, theta = (n_dk+a)/(L+sum a) and
    ll = getLLRest(model + docs) - getLLRest(model) + getLLDocs(docs) by scipy."""
    from scipy.special import gammaln
    words, off, nw, _ = synth.workload("plain_k20", n_docs=60)
    ws, offs, V, order = corpus.prepare_arrays(words, off, nw)
    K = 9
    alpha = (np.arange(K) * 0.01 + 0.05).astype(np.float32)
    m = O.Model(K, V, alpha, 0.02, 3)
    m.set_optim(0, 0)
    m.set_corpus(ws, offs)
    m.train(5, workers=2)
    _, _, nkw0, nk0 = m.state()
    uw, uoff = synth.unseen_docs(8, 12, V, 20, np.log(60.0), 1.0, 5, 300, zipf_s=None)
    uw, uoff = corpus.csr_from_arrays(uw, uoff, np.arange(V))
    for scheme, workers in [(1, 1), (3, 3), (2, 2)]:
        theta, ll, z = m.infer_together(uw, uoff, 7, infer_seed=11, workers=workers, scheme=scheme)
        D = len(uoff) - 1
        ndk = np.zeros((D, K), np.int64)
        nkw = nkw0.astype(np.int64).copy()
        for d in range(D):
            for g in range(int(uoff[d]), int(uoff[d + 1])):
                if uw[g] < V:
                    ndk[d, z[g]] += 1
                    nkw[uw[g], z[g]] += 1
        nk = nkw.sum(axis=0)
        assert np.array_equal(nk0 + ndk.sum(axis=0), nk)
        eta = float(np.float32(0.02))
        a = alpha.astype(np.float64)

        def rest(nkw_, nk_):
            return K * gammaln(V * eta) - np.sum(gammaln(nk_ + V * eta)) + \
                np.sum(np.where(nkw_ != 0, gammaln(nkw_ + eta) - gammaln(eta), 0.0))
        L = ndk.sum(axis=1)
        docs = np.sum(gammaln(a.sum()) - gammaln(L + a.sum())) + np.sum(gammaln(ndk + a) - gammaln(a))
        ref = rest(nkw, nk) - rest(nkw0.astype(np.int64), nk0.astype(np.int64)) + docs
        assert ll == pytest.approx(ref, rel=1e-9)
        asum = np.float32(0)
        for x in alpha:
            asum = np.float32(asum + x)
        th = ((ndk.astype(np.float32) + alpha) / (L.astype(np.float32)[:, None] + asum)).astype(np.float32)
        assert np.array_equal(theta, th)
