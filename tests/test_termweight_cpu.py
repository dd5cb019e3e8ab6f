"""TermWeight IDF / PMI on the CPU: the token weights and the oracle's
term-weighted sampler (float counts, clamped removal, weighted merge),
pinned by reductions to the integer sampler, by a plain-Python restatement
on tiny inputs and by closed forms of the log-likelihood."""
import math

import numpy as np
import pytest
from scipy.special import gammaln

from oracle import oracle as O
from tomotopy_amd import corpus, synth


def _corpus(n_docs, seed=0, min_cf=0, rm_top=0):
    words, off, nw, _ = synth.workload("plain_k20", n_docs=n_docs)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw, min_cf=min_cf, rm_top=rm_top)
    return ws, offs, V


@pytest.mark.parametrize("scheme", ["idf", "pmi"])
@pytest.mark.parametrize("min_cf,rm_top", [(0, 0), (3, 2)])
def test_term_weights_match_loop_restatement(scheme, min_cf, rm_top):
    ws, offs, V = _corpus(60, min_cf=min_cf, rm_top=rm_top)
    a = corpus.term_weights(ws, offs, V, scheme)
    b = O.term_weights(ws, offs, V, scheme)
    assert a.dtype == np.float32
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.all(a[ws >= V] == 0.0)


def test_term_weights_known_answers():
    # 3 documents over words 0..2 (sorted): word 0 in every document, word 1 in one, word 2 removed (V = 2)
    words = np.array([0, 1, 1, 2, 0, 0, 2], np.uint32)
    off = np.array([0, 4, 5, 7], np.uint64)
    idf = corpus.term_weights(words, off, 2, "idf")
    assert idf[0] == 0.0 and idf[4] == 0.0  # log(3/3)
    assert idf[1] == np.float32(math.log(3.0)) and idf[2] == idf[1]
    assert idf[3] == 0.0 and idf[6] == 0.0  # removed words
    pmi = corpus.term_weights(words, off, 2, "pmi")
    # totCf = 7 tokens; cf0 = 3, cf1 = 2.  doc 0: tf1 = 2, len 4 -> log((2 / (2/7)) / 4) = log(7/4)
    assert abs(float(pmi[1]) - math.log(7.0 / 4.0)) < 1e-6 and pmi[2] == pmi[1]
    # doc 0, word 0: log((1 / (3/7)) / 4) = log(7/12) < 0 -> 0
    assert pmi[0] == 0.0
    # doc 1 = [0]: log((1 / (3/7)) / 1) = log(7/3); doc 2 = [0, 2]: log((1 / (3/7)) / 2) = log(7/6)
    assert abs(float(pmi[4]) - math.log(7.0 / 3.0)) < 1e-6
    assert abs(float(pmi[5]) - math.log(7.0 / 6.0)) < 1e-6


@pytest.mark.parametrize("workers,scheme", [(1, 1), (3, 3), (3, 2), (5, 2)])
def test_unit_weights_reduce_to_integer_sampler(workers, scheme):
    """Weights of 1.0: float counts hold integers exactly and no removal
    clamps, so the trajectory, counts, alpha and log-likelihood equal the
    integer sampler's, under each schedule (for copy_merge the float merge
    local_0 + sum (local_w - old) is exact on integers and equals the integer
    old + sum (local_w - old))."""
    ws, offs, V = _corpus(50)
    K = 12
    mi = O.Model(K, V, 0.1, 0.01, 7)
    mw = O.WeightedModel(K, V, 0.1, 0.01, 7)
    for m in (mi, mw):
        m.set_optim(2, 0)
    mi.set_corpus(ws, offs)
    mw.set_corpus(ws, offs, np.ones(len(ws), np.float32))
    mi.train(5, workers=workers, scheme=scheme)
    mw.train(5, workers=workers, scheme=scheme)
    zi, ndki, nkwi, nki = mi.state()
    zw, ndkw, nkww, nkw_, docw = mw.state()
    assert np.array_equal(zi, zw)
    assert np.array_equal(ndki.astype(np.float32), ndkw)
    assert np.array_equal(nkwi.astype(np.float32), nkww)
    assert np.array_equal(nki.astype(np.float32), nkw_)
    assert np.array_equal(mi.alpha(), mw.alpha())
    assert mi.loglik(0) == pytest.approx(mw.loglik(0), rel=1e-12)
    assert np.array_equal(docw, np.diff(offs.astype(np.int64)).astype(np.float32) * 0 +
                          np.array([np.count_nonzero(ws[int(offs[d]):int(offs[d + 1])] < V)
                                    for d in range(len(offs) - 1)], np.float32))


@pytest.mark.parametrize("scheme", [3, 2])
def test_weight_two_with_doubled_priors_keeps_trajectory(scheme):
    """w = 2, alpha' = 2 alpha, eta' = 2 eta scales every topic weight by exactly
    2 (powers of two are exact), so the sampled topics are the integer
    sampler's; the counts are twice the integer ones."""
    ws, offs, V = _corpus(40)
    K = 9
    mi = O.Model(K, V, 0.1, 0.01, 3)
    mw = O.WeightedModel(K, V, 0.2, 0.02, 3)
    for m in (mi, mw):
        m.set_optim(0, 0)
    mi.set_corpus(ws, offs)
    mw.set_corpus(ws, offs, np.full(len(ws), 2.0, np.float32))
    mi.train(4, workers=2, scheme=scheme)
    mw.train(4, workers=2, scheme=scheme)
    zi, ndki, nkwi, nki = mi.state()
    zw, ndkw, nkww, nkw_, _ = mw.state()
    assert np.array_equal(zi, zw)
    assert np.array_equal(2 * nkwi.astype(np.float32), nkww)
    assert np.array_equal(2 * nki.astype(np.float32), nkw_)


def _py_weighted_sweeps(words, off, wts, K, V, alpha, eta, seed, sweeps, z0):
    """Sequential (none-scheme) weighted Gibbs in plain Python over float32
    scalars: tomoto's CntUpdater (removal max(x - w, 0)), the weights
    (n_dk+a)(n_kw+eta)/(n_k+V eta), then the device-order prefix sum and search
    (the oracle's scan and search, pinned by their own tests)."""
    f = np.float32
    D = len(off) - 1
    ndk = [[f(0)] * K for _ in range(D)]
    nkw = [[f(0)] * K for _ in range(V)]
    nk = [f(0)] * K
    z = list(int(x) for x in z0)
    for d in range(D):
        for g in range(int(off[d]), int(off[d + 1])):
            v = int(words[g])
            if v < V:
                k, w = z[g], f(wts[g])
                ndk[d][k] = f(ndk[d][k] + w); nkw[v][k] = f(nkw[v][k] + w); nk[k] = f(nk[k] + w)
    veta = f(f(eta) * f(V))
    hits = [0]

    def clamp(x):
        if x < 0:
            hits[0] += 1
            return f(0)
        return x
    for s in range(sweeps):
        # W = 1: every document in the forShuffled order of worker 0, round 0
        sseed = int(O.philox([0, 0, s, O.tag("shuffle")], [seed & 0xFFFFFFFF, seed >> 32])[0])
        order = O.shuffled_order(D, sseed)
        for d in order:
            for g in range(int(off[d]), int(off[d + 1])):
                v = int(words[g])
                if v >= V:
                    continue
                w = f(wts[g])
                k = z[g]
                ndk[d][k] = clamp(f(ndk[d][k] - w)); nkw[v][k] = clamp(f(nkw[v][k] - w)); nk[k] = clamp(f(nk[k] - w))
                p = [f(f(f(ndk[d][j] + f(alpha[j])) * f(nkw[v][j] + f(eta))) / f(nk[j] + veta)) for j in range(K)]
                cdf = O.prefix_sum_device(np.array(p, np.float32), O.device_kpl(K))
                k = O.search(cdf, K, O.token_uniform(seed, g, s, "sweep"))
                z[g] = k
                ndk[d][k] = f(ndk[d][k] + w); nkw[v][k] = f(nkw[v][k] + w); nk[k] = f(nk[k] + w)
        # weighted merge: n_k = float row sums over words, blocks of 256 words
        # summed in word order, then the block sums in block order
        nk = [f(0)] * K
        for v0 in range(0, V, 256):
            part = [f(0)] * K
            for v in range(v0, min(V, v0 + 256)):
                for j in range(K):
                    nkw[v][j] = clamp(nkw[v][j])
                    part[j] = f(part[j] + nkw[v][j])
            nk = [f(nk[j] + part[j]) for j in range(K)]
    return (np.array(z, np.uint16), np.array(ndk, np.float32), np.array(nkw, np.float32), np.array(nk, np.float32),
            hits[0])


@pytest.mark.parametrize("scheme", ["idf", "pmi", "ragged"])
def test_weighted_oracle_matches_python_restatement(scheme):
    ws, offs, V = _corpus(12)
    if scheme == "ragged":  # weights of mixed magnitudes make removals clamp at 0
        rng = np.random.default_rng(5)
        wts = (rng.random(len(ws)) * 3.0 + 1e-3).astype(np.float32)
    else:
        wts = corpus.term_weights(ws, offs, V, scheme)
    K = 7
    alpha = np.full(K, 0.1, np.float32)
    m = O.WeightedModel(K, V, alpha, 0.01, 21)
    m.set_optim(0, 0)
    m.set_corpus(ws, offs, wts)
    z0 = m.state()[0]
    m.train(3, workers=1)
    z, ndk, nkw, nk, _ = m.state()
    pz, pndk, pnkw, pnk, clamps = _py_weighted_sweeps(ws, offs, wts, K, V, alpha, 0.01, 21, 3, z0)
    print("removals clamped at 0:", clamps)
    if scheme == "ragged":
        assert clamps > 0
    assert np.array_equal(z, pz)
    assert np.array_equal(ndk, pndk)
    assert np.array_equal(nkw, pnkw)
    assert np.array_equal(nk, pnk)


def test_weighted_loglik_closed_form():
    ws, offs, V = _corpus(30)
    wts = corpus.term_weights(ws, offs, V, "pmi")
    K = 8
    alpha = (np.arange(K) * 0.03 + 0.05).astype(np.float32)
    m = O.WeightedModel(K, V, alpha, 0.02, 4)
    m.set_optim(0, 0)
    m.set_corpus(ws, offs, wts)
    m.train(2)
    _, ndk, nkw, nk, docw = m.state()
    a = alpha.astype(np.float64)
    docs = np.sum(gammaln(a.sum()) - gammaln(docw.astype(np.float64) + a.sum()))
    docs += np.sum(np.where(ndk != 0, gammaln(ndk.astype(np.float64) + a) - gammaln(a), 0.0))
    eta = float(np.float32(0.02))
    rest = K * gammaln(V * eta) - np.sum(gammaln(nk.astype(np.float64) + V * eta))
    rest += np.sum(np.where(nkw != 0, gammaln(nkw.astype(np.float64) + eta) - gammaln(eta), 0.0))
    ld, lr = m.loglik(0)
    assert ld == pytest.approx(docs, rel=1e-11)
    assert lr == pytest.approx(rest, rel=1e-11)
    # document weight sums are the sequential float sums of the weights
    for d in range(len(offs) - 1):
        acc = np.float32(0)
        for g in range(int(offs[d]), int(offs[d + 1])):
            if ws[g] < V:
                acc = np.float32(acc + wts[g])
        assert docw[d] == acc


def test_weighted_copy_merge_counts_follow_z():
    """copy_merge on IDF weights: after every merge n_kw equals the sum of the
    weights of the word's tokens on each topic (to float rounding: the counts
    are accumulated in float), n_k is the row sum of n_kw, and n_dk the weight
    sum of the document's tokens; a merge that dropped a worker's changes or
    added old twice breaks the first check by whole token weights."""
    ws, offs, V = _corpus(60, min_cf=2)
    wts = corpus.term_weights(ws, offs, V, "idf")
    K = 10
    m = O.WeightedModel(K, V, 0.1, 0.01, 11)
    m.set_optim(0, 0)
    m.set_corpus(ws, offs, wts)
    for _ in range(3):
        m.train(1, workers=4, scheme=2)
        z, ndk, nkw, nk, _ = m.state()
        exp = np.zeros((V, K), np.float64)
        valid = ws < V
        np.add.at(exp, (ws[valid].astype(np.int64), z[valid].astype(np.int64)), wts[valid].astype(np.float64))
        np.testing.assert_allclose(nkw.reshape(V, K), exp, rtol=0, atol=2e-4 * max(1.0, exp.max()))
        assert (nkw >= 0).all()
        np.testing.assert_allclose(nk, nkw.reshape(V, K).sum(axis=0, dtype=np.float64), rtol=1e-5)
