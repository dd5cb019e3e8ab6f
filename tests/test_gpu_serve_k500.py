"""The serving workload at its full size: a device-resident K=500, V=100,000
model (trained on 20,000 documents), requests of 1..64 documents with
lengths up to 10,000 tokens, 50 fold-in iterations (LDAModel::_infer,
LDAModel.hpp:920-953).  Every served result equals the stand-alone batched
fold-in (lda_infer mode 0) bit for bit, with exact sampling (no
segmentation) and with the segmentation the benchmark serves with; three
requests are checked against the CPU oracle's fold-in directly (theta bit
for bit, FP64 ll to 1e-9)."""
import numpy as np
import pytest

from oracle import oracle as O
from tomotopy_amd import _native, corpus, synth

pytestmark = pytest.mark.gpu

ITERS = 50


@pytest.fixture(scope="module")
def served_model():
    (tw, toff, nw), reqs_raw = synth.serve_workload(600)
    ws, offs, V, order = corpus.prepare_arrays(tw, toff, nw)
    reqs = [corpus.csr_from_arrays(w, o, order) for w, o in reqs_raw]
    ctx = _native.Context(500, V, 0.1, 0.01, seed=55)
    ctx.upload_corpus(ws, offs)
    ctx.train(50, workers=0)
    longest = [int(np.diff(o).max()) if len(o) > 1 else 0 for _, o in reqs]
    order_by_len = np.argsort(longest)[::-1]
    rng = np.random.default_rng(7)
    pick = list(order_by_len[:3]) + [int(i) for i in rng.choice(len(reqs), 21, replace=False)]
    picked = [reqs[i] for i in dict.fromkeys(pick)]
    print("V", V, "training tokens", len(ws), "requests", len(picked), "longest document",
          max(int(np.diff(o).max()) for _, o in picked), flush=True)
    return ctx, ws, offs, V, picked


@pytest.mark.parametrize("seg_tokens", [0, 128])
def test_serve_k500_equals_batched_infer(served_model, seg_tokens):
    ctx, _, _, V, reqs = served_model
    assert max(int(np.diff(o).max()) for _, o in reqs) >= 2000
    seg_max = ctx.infer_options()["seg_supported"] if seg_tokens else 1
    ctx.set_infer_options(seg_tokens, seg_max)
    try:
        seed = 2024
        with _native.Server(ctx, hbm_budget_bytes=1 << 30, iterations=ITERS, step_iters=ITERS, step_tokens=12000,
                            infer_seed=seed) as srv:
            ids = [srv.submit(uw, uoff) for uw, uoff in reqs]
            served = [srv.wait(i, timeout_ms=120000)[:2] for i in ids]
            st = srv.stats()
        print("seg_tokens", seg_tokens, "steps", st["steps"], "docs", st["docs_done"], flush=True)
        for (uw, uoff), (th, ll) in zip(reqs, served):
            th0, ll0, _ = ctx.infer(uw, uoff, ITERS, infer_seed=seed, mode=0)
            assert np.array_equal(th, th0)
            assert np.array_equal(ll, ll0)
    finally:
        ctx.set_infer_options(0, 1)


def test_serve_k500_matches_oracle(served_model):
    ctx, ws, offs, V, reqs = served_model
    z, _, _, _ = ctx.download_state()
    m = O.Model(500, V, ctx.get_alpha(), 0.01, 55, scan_order=1)
    m.set_corpus(ws, offs, z)
    small = sorted(range(len(reqs)), key=lambda i: len(reqs[i][0]))[:3]
    with _native.Server(ctx, hbm_budget_bytes=1 << 30, iterations=ITERS, step_iters=20, step_tokens=12000,
                        infer_seed=0) as srv:
        ids = [srv.submit(*reqs[i]) for i in small]
        for rid, i in zip(ids, small):
            th, ll, _ = srv.wait(rid, timeout_ms=120000)
            uw, uoff = reqs[i]
            print("request", i, "docs", len(uoff) - 1, "tokens", len(uw), flush=True)
            th_o, ll_o, _ = m.infer(uw, uoff, ITERS, infer_seed=0, mode=0)
            assert np.array_equal(th, th_o)
            np.testing.assert_allclose(ll, ll_o, rtol=1e-9)
