"""Word priors (eta_kv) on the device against the CPU oracle: the initial
draw from the eta columns, bit-exact none / partition / copy_merge schedules
(integer and term-weighted counts), the FP64 log-likelihood, the device-wide
sweep's invariants, and the LDAModel surface."""
import os

import numpy as np
import pytest

from oracle import oracle as O
import tomotopy_amd as tp
from tomotopy_amd import _native, corpus, synth

pytestmark = pytest.mark.gpu


def _setup(K, seed, n_docs=80, n_prior=25, tw=0, eta=0.01):
    words, off, nw, _ = synth.workload("plain_k20", n_docs=n_docs)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw)
    rng = np.random.default_rng(seed)
    eta_kv = np.full((V, K), np.float32(eta), np.float32)
    pw = rng.choice(V, size=min(n_prior, V), replace=False).astype(np.uint32)
    for v in pw:
        eta_kv[v] = (rng.random(K) * rng.choice([0.01, 1.0, 5.0])).astype(np.float32)
    ctx = _native.Context(K, V, 0.1, eta, seed=seed, tw=tw)
    ctx.set_word_priors(pw, eta_kv[pw])
    wts = None
    if tw:
        wts = corpus.term_weights(ws, offs, V, "pmi" if tw == 2 else "idf")
        m = O.WeightedModel(K, V, 0.1, eta, seed, scan_order=1)
    else:
        m = O.Model(K, V, 0.1, eta, seed, scan_order=1)
    m.set_priors(eta_kv)
    ctx.upload_corpus(ws, offs, weights=wts)
    if tw:
        m.set_corpus(ws, offs, wts)
    else:
        m.set_corpus(ws, offs)
    return ctx, m, ws, offs, V


def _same_int(ctx, m):
    zg, ndkg, nkwg, nkg = ctx.download_state()
    zo, ndko, nkwo, nko = m.state()
    assert np.array_equal(zg, zo), "z differs at %d tokens" % int((zg != zo).sum())
    assert np.array_equal(ndkg, ndko) and np.array_equal(nkwg, nkwo) and np.array_equal(nkg, nko)


@pytest.mark.parametrize("K,workers,scheme", [(20, 1, 1), (20, 3, 3), (20, 2, 2), (200, 4, 3), (1000, 2, 3),
                                              (1000, 2, 2)])
def test_prior_schedules_bit_exact(K, workers, scheme):
    ctx, m, _, _, _ = _setup(K, seed=K + workers)
    ctx.set_optim(0, 0)
    m.set_optim(0, 0)
    _same_int(ctx, m)  # initial draw from the eta columns
    for it in range(3):
        print("K", K, "workers", workers, "scheme", scheme, "sweep", it, flush=True)
        ctx.train(1, workers=workers, scheme=scheme)
        m.train(1, workers=workers, scheme=scheme)
        _same_int(ctx, m)
    gd, gr = ctx.loglik()
    od, orr = m.loglik(0)
    assert abs(gd - od) <= 1e-9 * abs(od) and abs(gr - orr) <= 1e-9 * abs(orr)


@pytest.mark.parametrize("K,scheme", [(20, 3), (200, 3), (20, 2), (200, 2)])
def test_prior_weighted_partition_bit_exact(K, scheme):
    ctx, m, _, _, _ = _setup(K, seed=7, tw=2)
    ctx.set_optim(0, 0)
    m.set_optim(0, 0)
    for it in range(3):
        print("K", K, "scheme", scheme, "sweep", it, flush=True)
        ctx.train(1, workers=3, scheme=scheme)
        m.train(1, workers=3, scheme=scheme)
        zg, ndkg, nkwg, nkg = ctx.download_state_f()
        zo, ndko, nkwo, nko, _ = m.state()
        assert np.array_equal(zg, zo)
        assert np.array_equal(nkwg.view(np.uint32), nkwo.view(np.uint32))
        assert np.array_equal(nkg.view(np.uint32), nko.view(np.uint32))
    gd, gr = ctx.loglik()
    od, orr = m.loglik(0)
    assert abs(gr - orr) <= 1e-9 * abs(orr)


@pytest.mark.parametrize("K", [20, 1000])
def test_prior_device_wide(K):
    ctx, m, ws, offs, V = _setup(K, seed=9, n_docs=300)
    ctx.set_optim(0, 0)
    m.set_optim(0, 0)
    ctx.train(30, workers=0)
    m.train(30, workers=1, scheme=1)
    assert not ctx.check_invariants().any()
    gd, gr = ctx.loglik()
    od, orr = m.loglik(0)
    n = int(np.count_nonzero(ws < V))
    print("ll_per_word device-wide %.5f oracle %.5f" % ((gd + gr) / n, (od + orr) / n))
    assert abs((gd + gr) - (od + orr)) <= 0.02 * abs(od + orr)


def test_model_surface_with_priors(tmp_path):
    sample = os.path.join(os.path.dirname(__file__), "golden", "sample.txt")
    mdl = tp.LDAModel(k=5, min_cf=3, seed=1)
    with open(sample, encoding="utf-8") as f:
        for i, line in enumerate(f):
            if i >= 300:
                break
            if line.split():
                mdl.add_doc(line.split())
    common = sorted(mdl.vocabs, key=lambda w: -mdl.vocab_freq[mdl.vocabs.index(w)])[:4]
    for w in common:
        mdl.set_word_prior(w, {2: 5.0, 0: 1e-4, 1: 1e-4, 3: 1e-4, 4: 1e-4})
    mdl.train(30, workers=1)
    top2 = [w for w, _ in mdl.get_topic_words(2, top_n=10)]
    assert all(w in top2 for w in common)
    assert mdl.get_word_prior(common[0]) == pytest.approx([1e-4, 1e-4, 5.0, 1e-4, 1e-4])
    path = str(tmp_path / "p.bin")
    mdl.save(path)
    m2 = tp.LDAModel.load(path)
    assert m2.get_word_prior(common[0]) == mdl.get_word_prior(common[0])
    assert m2.ll_per_word == pytest.approx(mdl.ll_per_word, rel=1e-12)
    m2.train(5, workers=1)
    dist, ll = m2.infer(m2.make_doc(common * 5), iterations=50)
    assert abs(sum(dist) - 1.0) < 1e-5 and np.isfinite(ll)
    assert int(np.argmax(dist)) == 2  # the words' prior holds them on topic 2
    dists, ll_all = m2.infer([m2.make_doc(common * 3), m2.make_doc(common)], together=True, iterations=20)
    assert len(dists) == 2 and np.isfinite(ll_all)
