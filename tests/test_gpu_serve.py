"""Continuous-batching fold-in server: every request's result equals the
stand-alone fold-in (lda_infer, itself pinned to the oracle) bit for bit,
whatever it was batched with and however its iterations were split into steps."""
import threading
import time

import numpy as np
import pytest

from oracle import oracle as O
from tomotopy_amd import _native, corpus, synth

pytestmark = pytest.mark.gpu


def _model(K, seed, n_docs=120, iters=10):
    words, off, nw, _ = synth.workload("plain_k20", n_docs=n_docs)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw)
    ctx = _native.Context(K, V, 0.1, 0.01, seed=seed)
    ctx.upload_corpus(ws, offs)
    ctx.train(iters, workers=0)
    return ctx, ws, offs, V


def _requests(V, n_req, seed, max_docs=12, lo=1, hi=900):
    rng = np.random.default_rng(seed)
    out = []
    for r in range(n_req):
        D = int(rng.integers(1, max_docs + 1))
        uw, uoff = synth.unseen_docs(seed * 1000 + r, D, V, 20, np.log(60.0), 1.2, lo, hi, zipf_s=None)
        uw, uoff = corpus.csr_from_arrays(uw, uoff, np.arange(V))
        out.append((uw, uoff))
    return out


def _check_same(ctx, reqs, results, iters, seed):
    for (uw, uoff), (th, ll) in zip(reqs, results):
        th0, ll0, _ = ctx.infer(uw, uoff, iters, infer_seed=seed, mode=0)
        assert np.array_equal(th, th0)
        assert np.array_equal(ll, ll0)


@pytest.mark.parametrize("K,step_iters,step_tokens", [(20, 50, 1 << 30), (20, 3, 1 << 30), (50, 50, 200),
                                                      (500, 7, 1500)])
def test_serve_equals_standalone_infer(K, step_iters, step_tokens):
    """All requests in flight at once; iterations split into steps of varying size."""
    ctx, _, _, V = _model(K, seed=3)
    iters, seed = 25, 99
    reqs = _requests(V, 24, seed=K + step_iters)
    print("K", K, "step_iters", step_iters, "step_tokens", step_tokens, flush=True)
    with _native.Server(ctx, hbm_budget_bytes=64 << 20, iterations=iters, step_iters=step_iters,
                        step_tokens=step_tokens, infer_seed=seed) as srv:
        ids = [srv.submit(uw, uoff) for uw, uoff in reqs]
        results = []
        for rid in ids:
            th, ll, lat = srv.wait(rid, timeout_ms=60000)
            assert lat >= 0.0
            results.append((th, ll))
        st = srv.stats()
    print("steps", st["steps"], "docs", st["docs_done"], flush=True)
    assert st["requests_done"] == len(reqs)
    assert st["docs_done"] == sum(len(o) - 1 for _, o in reqs)
    if step_iters < iters or step_tokens < 1 << 20:
        assert st["steps"] > 1
    _check_same(ctx, reqs, results, iters, seed)


def test_serve_matches_oracle():
    """The server's results against the CPU oracle's fold-in directly."""
    ctx, ws, offs, V = _model(20, seed=4, n_docs=100)
    z, _, _, _ = ctx.download_state()
    m = O.Model(20, V, ctx.get_alpha(), 0.01, 4, scan_order=1)
    m.set_corpus(ws, offs, z)
    reqs = _requests(V, 5, seed=17)
    with _native.Server(ctx, hbm_budget_bytes=16 << 20, iterations=12, step_iters=5, step_tokens=400,
                        infer_seed=0) as srv:
        ids = [srv.submit(uw, uoff) for uw, uoff in reqs]
        for rid, (uw, uoff) in zip(ids, reqs):
            th, ll, _ = srv.wait(rid, timeout_ms=60000)
            th_o, ll_o, _ = m.infer(uw, uoff, 12, infer_seed=0, mode=0)
            assert np.array_equal(th, th_o)
            np.testing.assert_allclose(ll, ll_o, rtol=1e-9)


def test_serve_arrivals_during_steps_and_admission_by_memory():
    """Requests arrive from another thread while earlier ones run; a 128 KiB budget
    admits only part of them at a time.  Results are unchanged."""
    ctx, _, _, V = _model(20, seed=5)
    iters, seed = 30, 7
    reqs = _requests(V, 60, seed=23, max_docs=8, hi=1500)
    results = [None] * len(reqs)
    with _native.Server(ctx, hbm_budget_bytes=128 << 10, iterations=iters, step_iters=4, step_tokens=2000,
                        infer_seed=seed) as srv:
        arena = srv.stats()["arena_tokens"]
        print("arena tokens", arena, flush=True)
        assert sum(len(w) for w, _ in reqs) > arena  # more work than fits at once
        ids = [None] * len(reqs)

        def client():
            rng = np.random.default_rng(1)
            for i, (uw, uoff) in enumerate(reqs):
                ids[i] = srv.submit(uw, uoff)
                time.sleep(float(rng.exponential(0.002)))

        t = threading.Thread(target=client)
        t.start()
        t.join()
        for i in range(len(reqs)):
            th, ll, _ = srv.wait(ids[i], timeout_ms=120000)
            results[i] = (th, ll)
        st = srv.stats()
    assert st["requests_done"] == len(reqs)
    _check_same(ctx, reqs, results, iters, seed)


def test_serve_errors_and_empty_requests():
    ctx, _, _, V = _model(20, seed=6, n_docs=40, iters=2)
    with _native.Server(ctx, hbm_budget_bytes=1 << 20, iterations=5, infer_seed=1) as srv:
        rid = srv.submit(np.zeros(0, np.uint32), np.zeros(1, np.uint64))
        th, ll, _ = srv.wait(rid, timeout_ms=1000)
        assert th.shape == (0, 20) and ll.shape == (0,)
        # a request with an empty document and a normal one
        rid = srv.submit(np.array([1, 2, 2], np.uint32), np.array([0, 0, 3], np.uint64))
        th, ll, _ = srv.wait(rid, timeout_ms=30000)
        th0, ll0, _ = ctx.infer(np.array([1, 2, 2], np.uint32), np.array([0, 0, 3], np.uint64), 5, infer_seed=1)
        assert np.array_equal(th, th0) and np.array_equal(ll, ll0)
        with pytest.raises(ValueError):
            srv.submit(np.array([3, 1], np.uint32), np.array([0, 2], np.uint64))
        too_many = int(srv.stats()["arena_tokens"]) + 1
        with pytest.raises(MemoryError):
            srv.submit(np.zeros(too_many, np.uint32), np.array([0, too_many], np.uint64))
        with pytest.raises(ValueError):
            srv.wait(123456)


@pytest.mark.parametrize("K,step_iters,step_tokens", [(20, 4, 300), (200, 50, 1 << 30)])
def test_serve_with_word_priors(K, step_iters, step_tokens):
    """A model with word priors: the served results equal the stand-alone
    fold-in bit for bit, and that one equals the oracle's fold-in with eta_kv
    (initial topics from the eta columns)."""
    words, off, nw, _ = synth.workload("plain_k20", n_docs=120)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw)
    rng = np.random.default_rng(K)
    eta_kv = np.full((V, K), np.float32(0.01), np.float32)
    pw = rng.choice(V, size=30, replace=False).astype(np.uint32)
    for v in pw:
        eta_kv[v] = (rng.random(K) * rng.choice([0.01, 1.0, 5.0])).astype(np.float32)
    ctx = _native.Context(K, V, 0.1, 0.01, seed=K)
    ctx.set_word_priors(pw, eta_kv[pw])
    ctx.upload_corpus(ws, offs)
    ctx.train(8, workers=0)
    z, _, _, _ = ctx.download_state()
    m = O.Model(K, V, ctx.get_alpha(), 0.01, K, scan_order=1)
    m.set_priors(eta_kv)
    m.set_corpus(ws, offs, z)
    iters, seed = 12, 5
    reqs = _requests(V, 10, seed=K + 1)
    print("priors K", K, flush=True)
    with _native.Server(ctx, hbm_budget_bytes=32 << 20, iterations=iters, step_iters=step_iters,
                        step_tokens=step_tokens, infer_seed=seed) as srv:
        ids = [srv.submit(uw, uoff) for uw, uoff in reqs]
        results = [srv.wait(rid, timeout_ms=60000)[:2] for rid in ids]
    _check_same(ctx, reqs, results, iters, seed)
    for (uw, uoff), (th, ll) in zip(reqs[:4], results[:4]):
        th_o, ll_o, _ = m.infer_gen(uw, uoff, iters, infer_seed=seed)
        assert np.array_equal(th, th_o)
        np.testing.assert_allclose(ll, ll_o, rtol=1e-9)
