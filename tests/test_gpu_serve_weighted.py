"""The server for TermWeight IDF / PMI models and for K above 2048 (the
run-column fold-in, each document folded in whole within one step): every
request's result equals the batched lda_infer_w / lda_infer (mode 0, itself
pinned to the oracle) bit for bit, alone or in any mix, on one or several
lanes; three requests are checked against the oracle directly."""
import threading

import numpy as np
import pytest

from oracle import oracle as O
from tomotopy_amd import _native, corpus, synth

pytestmark = pytest.mark.gpu

TW = {"one": 0, "idf": 1, "pmi": 2}


def _model(kind, K, seed, n_docs=100):
    words, off, nw, _ = synth.workload("plain_k20", n_docs=n_docs)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw, min_cf=2, rm_top=1)
    tw = TW[kind]
    wts = corpus.term_weights(ws, offs, V, kind) if tw else None
    ctx = _native.Context(K, V, 0.1, 0.01, seed=seed, tw=tw)
    ctx.upload_corpus(ws, offs, weights=wts)
    ctx.set_optim(0, 0)
    ctx.train(4, workers=1, scheme=1)
    m = (O.WeightedModel if tw else O.Model)(K, V, 0.1, 0.01, seed, scan_order=1)
    if tw:
        m.set_corpus(ws, offs, wts)
    else:
        m.set_corpus(ws, offs)
    m.set_optim(0, 0)
    m.train(4, workers=1, scheme=1)
    vocab_w = None
    if tw:
        # an unseen token's weight is its word's vocabulary weight (LDAModel.hpp:794-809)
        vocab_w = corpus.vocab_weights(ws, offs, V, kind)
    return ctx, m, V, vocab_w


def _requests(V, n_req, seed, vocab_w, max_docs=8):
    rng = np.random.default_rng(seed)
    out = []
    for r in range(n_req):
        D = int(rng.integers(1, max_docs + 1))
        uw, uoff = synth.unseen_docs(seed * 1000 + r, D, V, 20, np.log(60.0), 1.2, 1, 600, zipf_s=None)
        uw, uoff = corpus.csr_from_arrays(uw, uoff, np.arange(V))
        wt = vocab_w[uw].astype(np.float32) if vocab_w is not None else None
        out.append((uw, uoff, wt))
    return out


@pytest.mark.parametrize("kind,K,lanes", [("pmi", 20, 1), ("idf", 200, 2), ("pmi", 1000, 1), ("one", 3000, 2),
                                          ("pmi", 4096, 1)])
def test_serve_run_column_equals_batched(kind, K, lanes):
    ctx, m, V, vocab_w = _model(kind, K, seed=K + lanes)
    iters, seed = 12, 77
    reqs = _requests(V, 16, seed=K, vocab_w=vocab_w)
    print(kind, "K", K, "lanes", lanes, flush=True)
    with _native.Server(ctx, hbm_budget_bytes=256 << 20, iterations=iters, infer_seed=seed, lanes=lanes) as srv:
        ids = [srv.submit(uw, uoff, weights=wt) for uw, uoff, wt in reqs]
        results = [srv.wait(rid, timeout_ms=120000)[:2] for rid in ids]
        st = srv.stats()
    print("steps", st["steps"], "docs", st["docs_done"], flush=True)
    assert st["docs_done"] == sum(len(uoff) - 1 for _, uoff, _ in reqs)
    for (uw, uoff, wt), (th, ll) in zip(reqs, results):
        th0, ll0, _ = ctx.infer(uw, uoff, iters, infer_seed=seed, mode=0, weights=wt)
        assert np.array_equal(th, th0)
        assert np.array_equal(ll, ll0)
    # the first three requests against the oracle directly
    for (uw, uoff, wt), (th, ll) in list(zip(reqs, results))[:3]:
        if kind == "one":
            tho, llo, _ = m.infer(uw, uoff, iters, infer_seed=seed)
        else:
            tho, llo, _ = m.infer_gen(uw, uoff, iters, infer_seed=seed, weights=wt)
        assert np.array_equal(th, tho)
        np.testing.assert_allclose(ll, llo, rtol=1e-9, atol=1e-9)


def test_serve_weighted_concurrent_submitters_and_errors():
    """Requests from several threads; a request alone equals the same request
    in the mix; missing or bad weights are refused."""
    ctx, _, V, vocab_w = _model("pmi", 200, seed=5)
    iters, seed = 10, 3
    reqs = _requests(V, 24, seed=9, vocab_w=vocab_w)
    out = [None] * len(reqs)
    with _native.Server(ctx, hbm_budget_bytes=128 << 20, iterations=iters, infer_seed=seed, lanes=2) as srv:
        def run(lo):
            for i in range(lo, len(reqs), 3):
                uw, uoff, wt = reqs[i]
                out[i] = srv.wait(srv.submit(uw, uoff, weights=wt), timeout_ms=120000)[:2]
        th = [threading.Thread(target=run, args=(i,)) for i in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        uw, uoff, wt = reqs[5]
        alone = srv.wait(srv.submit(uw, uoff, weights=wt), timeout_ms=120000)[:2]
        assert np.array_equal(alone[0], out[5][0]) and np.array_equal(alone[1], out[5][1])
        with pytest.raises(ValueError):
            srv.submit(uw, uoff)  # a term-weighted model's server needs the weights
        bad = wt.copy()
        bad[0] = -1.0
        with pytest.raises(ValueError):
            srv.submit(uw, uoff, weights=bad)
    for (uw, uoff, wt), (th, ll) in zip(reqs, out):
        th0, ll0, _ = ctx.infer(uw, uoff, iters, infer_seed=seed, mode=0, weights=wt)
        assert np.array_equal(th, th0) and np.array_equal(ll, ll0)
