"""TermWeight IDF / PMI training on the device against the CPU oracle:
bit-exact under the deterministic schedules (none, partition, copy_merge), FP64
log-likelihood to 1e-9, device-wide sweep by its invariants and by the
per-word log-likelihood of an independent chain."""
import numpy as np
import pytest

from oracle import oracle as O
from tomotopy_amd import LDAModel, TermWeight, _native, corpus, synth

pytestmark = pytest.mark.gpu

TW = {"idf": 1, "pmi": 2}


def _corpus(n_docs, min_cf=0, rm_top=0):
    words, off, nw, _ = synth.workload("plain_k20", n_docs=n_docs)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw, min_cf=min_cf, rm_top=rm_top)
    return ws, offs, V


def _pair(scheme, K, seed, n_docs=80, alpha=0.1, eta=0.01, **kw):
    ws, offs, V = _corpus(n_docs, **kw)
    wts = corpus.term_weights(ws, offs, V, scheme)
    ctx = _native.Context(K, V, alpha, eta, seed=seed, tw=TW[scheme])
    ctx.upload_corpus(ws, offs, weights=wts)
    m = O.WeightedModel(K, V, alpha, eta, seed, scan_order=1)
    m.set_corpus(ws, offs, wts)
    return ctx, m, ws, offs, V, wts


def _compare(ctx, m):
    zg, ndkg, nkwg, nkg = ctx.download_state_f()
    zo, ndko, nkwo, nko, _ = m.state()
    assert np.array_equal(zg, zo), "z differs at %d tokens" % int((zg != zo).sum())
    assert np.array_equal(ndkg.view(np.uint32), ndko.view(np.uint32))
    assert np.array_equal(nkwg.view(np.uint32), nkwo.view(np.uint32))
    assert np.array_equal(nkg.view(np.uint32), nko.view(np.uint32))


@pytest.mark.parametrize("scheme", ["idf", "pmi"])
@pytest.mark.parametrize("K,workers", [(20, 1), (20, 3), (200, 4), (1000, 2)])
def test_weighted_schedule_bit_exact(scheme, K, workers):
    ctx, m, _, _, _, _ = _pair(scheme, K, seed=K + workers, min_cf=2, rm_top=1)
    ctx.set_optim(0, 0)
    m.set_optim(0, 0)
    _compare(ctx, m)  # initial float counts
    for it in range(3):
        print(scheme, "K", K, "workers", workers, "sweep", it, flush=True)
        ctx.train(1, workers=workers, scheme=3 if workers > 1 else 1)
        m.train(1, workers=workers, scheme=3 if workers > 1 else 1)
        _compare(ctx, m)


@pytest.mark.parametrize("scheme", ["idf", "pmi"])
@pytest.mark.parametrize("K,workers", [(20, 3), (200, 4), (1000, 2)])
def test_weighted_copy_merge_bit_exact(scheme, K, workers):
    ctx, m, _, _, _, _ = _pair(scheme, K, seed=K + 3 * workers, min_cf=2, rm_top=1)
    ctx.set_optim(0, 0)
    m.set_optim(0, 0)
    for it in range(3):
        print(scheme, "copy_merge K", K, "workers", workers, "sweep", it, flush=True)
        ctx.train(1, workers=workers, scheme=2)
        m.train(1, workers=workers, scheme=2)
        _compare(ctx, m)


@pytest.mark.parametrize("scheme", ["idf", "pmi"])
def test_weighted_alpha_optimisation_and_loglik(scheme):
    ctx, m, _, _, _, _ = _pair(scheme, 20, seed=5, n_docs=100)
    ctx.set_optim(4, 0)
    m.set_optim(4, 0)
    ctx.train(4, workers=2, scheme=3)
    m.train(4, workers=2, scheme=3)
    np.testing.assert_allclose(ctx.get_alpha(), m.alpha(), rtol=2e-6)
    _compare(ctx, m)
    m.set_alpha(ctx.get_alpha())
    gd, gr = ctx.loglik()
    od, orr = m.loglik(0)
    assert abs(gd - od) <= 1e-9 * abs(od)
    assert abs(gr - orr) <= 1e-9 * abs(orr)


@pytest.mark.parametrize("scheme,K", [("idf", 20), ("pmi", 200), ("pmi", 1000)])
def test_weighted_device_wide_sweep(scheme, K):
    """Device-wide schedule (float atomics, order not fixed): counts stay
    non-negative, n_k is the float row sum of n_kw, document rows sum to
    the document weights, and the per-word log-likelihood tracks an oracle
    chain of the none schedule within 2%."""
    ctx, m, ws, offs, V, wts = _pair(scheme, K, seed=8, n_docs=400)
    ctx.set_optim(0, 0)
    m.set_optim(0, 0)
    print(scheme, "K", K, flush=True)
    ctx.train(30, workers=0)
    m.train(30, workers=1, scheme=1)
    z, ndk, nkw, nk = ctx.download_state_f()
    valid = ws < V
    assert z[valid].max() < K
    assert (ndk >= 0).all() and (nkw >= 0).all() and (nk >= 0).all()
    # float row sums: blocks of 256 words in word order, then the blocks in order
    blocks = [np.cumsum(nkw[v0:v0 + 256], axis=0, dtype=np.float32)[-1] for v0 in range(0, V, 256)]
    seq = np.cumsum(np.array(blocks), axis=0, dtype=np.float32)[-1]
    assert np.array_equal(nk.view(np.uint32), seq.view(np.uint32))
    _, _, _, _, docw = m.state()
    np.testing.assert_allclose(ndk.sum(axis=1), docw, rtol=2e-3, atol=1e-3)
    wn = ctx.weighted_n()
    assert wn == pytest.approx(float(np.sum(wts[valid].astype(np.float64))), rel=1e-12)
    gd, gr = ctx.loglik()
    od, orr = m.loglik(0)
    llg, llo = (gd + gr) / wn, (od + orr) / wn
    print("ll_per_word device-wide %.5f oracle %.5f" % (llg, llo))
    assert abs(llg - llo) <= 0.02 * abs(llo)


@pytest.mark.parametrize("tw", [TermWeight.IDF, TermWeight.PMI])
def test_model_surface_term_weighted(tw):
    mdl = LDAModel(tw=tw, k=10, min_cf=2, rm_top=1, seed=42)
    rng = np.random.default_rng(0)
    vocab = ["w%d" % i for i in range(300)]
    for _ in range(150):
        n = int(rng.integers(5, 60))
        mdl.add_doc([vocab[int(i)] for i in rng.zipf(1.3, n) % 300])
    mdl.train(20, workers=0)
    assert mdl.tw == int(tw)
    llpw = mdl.ll_per_word
    assert np.isfinite(llpw)
    assert sum(mdl.get_count_by_topics()) == mdl.num_words
    dist = np.asarray(mdl.get_topic_word_dist(0), np.float64)
    assert abs(dist.sum() - 1.0) < 1e-3
    wf = np.asarray(mdl.used_vocab_weighted_freq)
    assert wf.shape == (len(mdl.used_vocabs),) and (wf >= 0).all()
    d0 = mdl.docs[0]
    assert np.isfinite(d0.get_ll())
    dist, ll = mdl.infer(mdl.make_doc(["w1", "w2", "w2"]))
    assert abs(sum(dist) - 1.0) < 1e-5 and np.isfinite(ll)
    dists, ll_all = mdl.infer([mdl.make_doc(["w1", "w2"]), mdl.make_doc(["w3", "w1"])], together=True)
    assert len(dists) == 2 and np.isfinite(ll_all)
