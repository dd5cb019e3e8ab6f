"""Word priors (set_word_prior, eta_kv) in the oracle on the CPU: the initial
draw (sampleFromDiscrete), the prior-weighted sampler against a plain-Python
restatement, the log-likelihood against scipy, and the LDAModel argument
rules."""
import numpy as np
import pytest
from scipy.special import gammaln

from oracle import oracle as O
from tomotopy_amd import LDAModel, corpus, synth


def _py_sample_from_discrete(w, u):
    """sample.hpp:121-131 in float32 scalars."""
    f = np.float32
    tot = f(0)
    for x in w:
        tot = f(tot + f(x))
    r = f(f(u) * tot)
    z = 0
    while r > f(w[z]) and z < len(w) - 1:
        r = f(r - f(w[z]))
        z += 1
    return z


def test_sample_from_discrete_known_answers_and_restatement():
    w = np.array([1.0, 2.0, 3.0], np.float32)  # total 6: r = 6u
    assert O.sample_from_discrete(w, 0.0) == 0          # r = 0 is not > 1
    assert O.sample_from_discrete(w, 1.0 / 6) == 0      # r = 1 is not > 1
    assert O.sample_from_discrete(w, 0.2) == 1          # r = 1.2 > 1, then 0.2 <= 2
    assert O.sample_from_discrete(w, 0.5) == 1          # r = 3 > 1, then 2 is not > 2
    assert O.sample_from_discrete(w, 0.51) == 2
    assert O.sample_from_discrete(w, 0.999999) == 2     # the last index takes the rest
    rng = np.random.default_rng(3)
    for _ in range(500):
        K = int(rng.integers(1, 40))
        w = (rng.random(K) * rng.choice([1e-3, 1.0, 50.0])).astype(np.float32)
        w[rng.random(K) < 0.2] = 0.0
        u = float(np.float32(rng.random()))
        assert O.sample_from_discrete(w, u) == _py_sample_from_discrete(w, u)


def _prior_table(V, K, eta, rng, n_prior):
    eta_kv = np.full((V, K), np.float32(eta), np.float32)
    words = rng.choice(V, size=n_prior, replace=False)
    for v in words:
        eta_kv[v] = (rng.random(K) * rng.choice([0.01, 1.0, 5.0])).astype(np.float32)
    return eta_kv


def _py_prior_sweeps(words, off, K, V, alpha, eta_kv, seed, sweeps, z0):
    """The none schedule with eta_kv in plain Python: weights
    (n_dk+a)(n_kw+eta_kv)/(n_k + sum_v eta_kv), the oracle's device-order
    scan and search (pinned by their own tests)."""
    f = np.float32
    D = len(off) - 1
    es = [f(0)] * K
    for v in range(V):
        for k in range(K):
            es[k] = f(es[k] + eta_kv[v, k])
    ndk = np.zeros((D, K), np.int64)
    nkw = np.zeros((V, K), np.int64)
    nk = np.zeros(K, np.int64)
    z = [int(x) for x in z0]
    for d in range(D):
        for g in range(int(off[d]), int(off[d + 1])):
            if words[g] < V:
                ndk[d, z[g]] += 1; nkw[words[g], z[g]] += 1; nk[z[g]] += 1
    kpl = O.device_kpl(K)
    for s in range(sweeps):
        sseed = int(O.philox([0, 0, s, O.tag("shuffle")], [seed & 0xFFFFFFFF, seed >> 32])[0])
        for d in O.shuffled_order(D, sseed):
            for g in range(int(off[d]), int(off[d + 1])):
                v = int(words[g])
                if v >= V:
                    continue
                k = z[g]
                ndk[d, k] -= 1; nkw[v, k] -= 1; nk[k] -= 1
                p = [f(f(f(f(ndk[d, j]) + f(alpha[j])) * f(f(nkw[v, j]) + eta_kv[v, j])) / f(f(nk[j]) + es[j]))
                     for j in range(K)]
                cdf = O.prefix_sum_device(np.array(p, np.float32), kpl)
                k = O.search(cdf, K, O.token_uniform(seed, g, s, "sweep"))
                z[g] = k
                ndk[d, k] += 1; nkw[v, k] += 1; nk[k] += 1
    return np.array(z, np.uint16), ndk, nkw, nk


def test_prior_sampler_matches_python_restatement():
    words, off, nw, _ = synth.workload("plain_k20", n_docs=14)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw)
    K, seed = 6, 19
    rng = np.random.default_rng(2)
    eta_kv = _prior_table(V, K, 0.01, rng, 12)
    alpha = np.full(K, 0.1, np.float32)
    m = O.Model(K, V, alpha, 0.01, seed)
    m.set_optim(0, 0)
    m.set_priors(eta_kv)
    m.set_corpus(ws, offs)
    z0 = m.state()[0]
    # the initial draw follows each word's eta column
    for g in range(len(ws)):
        if ws[g] < V:
            assert z0[g] == _py_sample_from_discrete(eta_kv[ws[g]], O.token_uniform(seed, g, 0, "init"))
    m.train(3, workers=1, scheme=1)
    z, ndk, nkw, nk = m.state()
    pz, pndk, pnkw, pnk = _py_prior_sweeps(ws, offs, K, V, alpha, eta_kv, seed, 3, z0)
    assert np.array_equal(z, pz)
    assert np.array_equal(ndk, pndk) and np.array_equal(nkw, pnkw) and np.array_equal(nk, pnk)


def test_prior_loglik_closed_form():
    words, off, nw, _ = synth.workload("plain_k20", n_docs=40)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw)
    K = 7
    rng = np.random.default_rng(4)
    eta_kv = _prior_table(V, K, 0.02, rng, 30)
    alpha = (np.arange(K) * 0.02 + 0.05).astype(np.float32)
    m = O.Model(K, V, alpha, 0.02, 8)
    m.set_optim(0, 0)
    m.set_priors(eta_kv)
    m.set_corpus(ws, offs)
    m.train(2, workers=2)
    _, ndk, nkw, nk = m.state()
    e = eta_kv.astype(np.float64)
    es = e.sum(axis=0)
    rest = np.sum(gammaln(es) - gammaln(nk + es))
    rest += np.sum(np.where(nkw != 0, gammaln(nkw + e) - gammaln(e), 0.0))
    ld, lr = m.loglik(0)
    assert lr == pytest.approx(rest, rel=1e-11)
    # a strong prior pulls its word to the favoured topic
    strong = np.full((V, K), np.float32(0.02), np.float32)
    top = np.argsort(-np.bincount(ws[ws < V], minlength=V))[:5]
    strong[top] = 0.0001
    strong[top, 3] = 50.0
    m2 = O.Model(K, V, alpha, 0.02, 8)
    m2.set_priors(strong)
    m2.set_corpus(ws, offs)
    m2.train(10, workers=1, scheme=1)
    _, _, nkw2, _ = m2.state()
    assert np.all(nkw2[top, 3] >= 0.9 * nkw2[top].sum(axis=1))


def test_set_word_prior_rules():
    mdl = LDAModel(k=3, eta=0.01)
    mdl.add_doc(["apple", "pie", "apple"])
    mdl.set_word_prior("apple", [0.01, 0.9, 0.01])
    assert mdl.get_word_prior("apple") == pytest.approx([0.01, 0.9, 0.01])
    mdl.set_word_prior("apple", {1: 0.5})  # the first prior of a word stays (emplace)
    assert mdl.get_word_prior("apple") == pytest.approx([0.01, 0.9, 0.01])
    mdl.set_word_prior("pear", {1: 0.9})
    assert mdl.get_word_prior("pear") == pytest.approx([0.01, 0.9, 0.01])
    assert mdl.get_word_prior("pie") == []
    with pytest.raises(ValueError):
        mdl.set_word_prior("x", [0.1, 0.2])
    with pytest.raises(ValueError):
        mdl.set_word_prior("x", [0.1, -0.2, 0.1])
    with pytest.raises(ValueError):
        mdl.set_word_prior("x", {5: 0.1})
    assert "pear" in mdl.vocabs  # setWordPrior adds the word to the dictionary
