"""Fold-in with together=True (LDAModel.hpp:852-892) on the device against
the CPU oracle: identical topics and theta under the deterministic schemes,
FP64 log-likelihood to 1e-9, and the LDAModel.infer surface."""
import os

import numpy as np
import pytest

from oracle import oracle as O
import tomotopy_amd as tp
from tomotopy_amd import _native, corpus, synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("K,workers,scheme", [(20, 1, 1), (20, 3, 3), (20, 2, 2), (200, 4, 3), (200, 2, 2)])
def test_together_bit_exact(K, workers, scheme):
    words, off, nw, _ = synth.workload("plain_k20", n_docs=80)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw)
    ctx = _native.Context(K, V, 0.1, 0.01, seed=K)
    ctx.upload_corpus(ws, offs)
    ctx.set_optim(0, 0)
    ctx.train(5, workers=0)
    z, _, _, _ = ctx.download_state()
    m = O.Model(K, V, ctx.get_alpha(), 0.01, K, scan_order=1)
    m.set_corpus(ws, offs, z)
    uw, uoff = synth.unseen_docs(K + scheme, 30, V, 20, np.log(80.0), 1.0, 5, 600, zipf_s=None)
    uw, uoff = corpus.csr_from_arrays(uw, uoff, np.arange(V))
    print("K", K, "workers", workers, "scheme", scheme, "docs", len(uoff) - 1, flush=True)
    th, ll, zg = ctx.infer_together(uw, uoff, 12, infer_seed=5, workers=workers, scheme=scheme)
    th_o, ll_o, zo = m.infer_together(uw, uoff, 12, infer_seed=5, workers=workers, scheme=scheme)
    assert np.array_equal(zg, zo), "z differs at %d tokens" % int((zg != zo).sum())
    assert np.array_equal(th, th_o)
    assert ll == pytest.approx(ll_o, rel=1e-9)


def test_infer_together_surface():
    sample = os.path.join(os.path.dirname(__file__), "golden", "sample.txt")
    mdl = tp.LDAModel(k=8, min_cf=2, seed=3)
    lines = [l.split() for l in open(sample, encoding="utf-8")]
    n_train = len(lines) * 4 // 5
    for ch in lines[:n_train]:
        if ch:
            mdl.add_doc(ch)
    mdl.train(20)
    docs = [mdl.make_doc(ch) for ch in lines[n_train:] if ch]
    assert docs
    assert mdl.infer([], together=True) == ([], 0.0)
    dists, ll = mdl.infer(docs, iterations=20, together=True)
    assert isinstance(ll, float) and np.isfinite(ll)
    assert len(dists) == len(docs) and all(abs(sum(d) - 1.0) < 1e-4 for d in dists)
    d0, ll0 = mdl.infer(docs[0], iterations=20, together=True, workers=1)
    assert len(d0) == 8 and np.isfinite(ll0)
    dists2, ll2 = mdl.infer(docs, iterations=20, together=True, workers=1, parallel=tp.ParallelScheme.PARTITION)
    dists3, ll3 = mdl.infer(docs, iterations=20, together=True, workers=1, parallel=tp.ParallelScheme.PARTITION)
    assert dists2 == dists3 and ll2 == ll3  # deterministic under a fixed scheme


def _gen_pair(kind, K, seed):
    """Device context and oracle model with the same trained state: an IDF /
    PMI model (float counts) and / or word priors."""
    words, off, nw, _ = synth.workload("plain_k20", n_docs=80)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw, min_cf=2)
    tw = {"prior": 0, "idf": 1, "pmi": 2, "pmi_prior": 2}[kind]
    wts = corpus.term_weights(ws, offs, V, "pmi" if tw == 2 else "idf") if tw else None
    ctx = _native.Context(K, V, 0.1, 0.01, seed=seed, tw=tw)
    m = (O.WeightedModel if tw else O.Model)(K, V, 0.1, 0.01, seed, scan_order=1)
    if kind.endswith("prior"):
        rng = np.random.default_rng(seed)
        eta_kv = np.full((V, K), np.float32(0.01), np.float32)
        pw = rng.choice(V, size=15, replace=False).astype(np.uint32)
        for v in pw:
            eta_kv[v] = (rng.random(K) * rng.choice([0.01, 1.0, 5.0])).astype(np.float32)
        ctx.set_word_priors(pw, eta_kv[pw])
        m.set_priors(eta_kv)
    ctx.upload_corpus(ws, offs, weights=wts)
    if tw:
        m.set_corpus(ws, offs, wts)
    else:
        m.set_corpus(ws, offs)
    for x in (ctx, m):
        x.set_optim(0, 0)
    ctx.train(3, workers=1, scheme=1)
    m.train(3, workers=1, scheme=1)
    vw = corpus.vocab_weights(ws, offs, V, "pmi" if tw == 2 else "idf") if tw else None
    return ctx, m, V, tw, vw


@pytest.mark.parametrize("kind,K,workers,scheme", [("prior", 20, 3, 3), ("prior", 200, 2, 2), ("idf", 20, 1, 1),
                                                   ("idf", 200, 3, 3), ("pmi", 20, 2, 2), ("pmi_prior", 200, 3, 3)])
def test_together_term_weighted_and_priors_bit_exact(kind, K, workers, scheme):
    ctx, m, V, tw, vw = _gen_pair(kind, K, seed=K + workers)
    uw, uoff = synth.unseen_docs(K + scheme, 25, V, 20, np.log(80.0), 1.0, 5, 600, zipf_s=None)
    uw, uoff = corpus.csr_from_arrays(uw, uoff, np.arange(V))
    uwt = corpus.term_weights(uw, uoff, V, "pmi" if tw == 2 else "idf", vocab_w=vw) if tw else None
    print(kind, "K", K, "workers", workers, "scheme", scheme, flush=True)
    th, ll, zg = ctx.infer_together(uw, uoff, 10, infer_seed=6, workers=workers, scheme=scheme, weights=uwt)
    th_o, ll_o, zo = m.infer_together(uw, uoff, 10, infer_seed=6, workers=workers, scheme=scheme, weights=uwt)
    assert np.array_equal(zg, zo), "z differs at %d tokens" % int((zg != zo).sum())
    assert np.array_equal(th, th_o)
    assert ll == pytest.approx(ll_o, rel=1e-9)
