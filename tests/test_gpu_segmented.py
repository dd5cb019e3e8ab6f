"""Segmented fold-in on the device against the CPU oracle's segmented rule
(bit-exact topics and theta, FP64 log-likelihood to 1e-9), in the batched
graph, the plain per-document path and the server."""
import numpy as np
import pytest

from oracle import oracle as O
from tomotopy_amd import _native, corpus, synth

pytestmark = pytest.mark.gpu


def _setup(K, seed, n_docs=120, iters=8):
    words, off, nw, _ = synth.workload("plain_k20", n_docs=n_docs)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw)
    ctx = _native.Context(K, V, 0.1, 0.01, seed=seed)
    ctx.upload_corpus(ws, offs)
    ctx.train(iters, workers=0)
    z, _, _, _ = ctx.download_state()
    m = O.Model(K, V, ctx.get_alpha(), 0.01, seed, scan_order=1)
    m.set_corpus(ws, offs, z)
    return ctx, m, V


def _docs(V, seed, n=30, mu=300.0, lo=2, hi=3000):
    uw, uoff = synth.unseen_docs(seed, n, V, 20, np.log(mu), 1.0, lo, hi, zipf_s=None)
    uw, uoff = corpus.csr_from_arrays(uw, uoff, np.arange(V))
    # a document with one word repeated (one long run) and one with words >= V
    rep = np.full(700, 3, np.uint32)
    bad = np.sort(np.concatenate([np.arange(0, 400, 2), np.full(50, V + 7)])).astype(np.uint32)
    lens = np.concatenate([np.diff(uoff.astype(np.int64)), [len(rep), len(bad)]])
    off = np.zeros(len(lens) + 1, np.uint64)
    off[1:] = np.cumsum(lens)
    return np.concatenate([uw, rep, bad]).astype(np.uint32), off


@pytest.mark.parametrize("K,seg_tokens,seg_max", [(20, 64, 8), (20, 200, 2), (200, 100, 4), (500, 128, 8),
                                                  (1000, 256, 8), (20, 1, 8)])
def test_segmented_infer_bit_exact(K, seg_tokens, seg_max):
    ctx, m, V = _setup(K, seed=K + seg_tokens)
    sup = ctx.infer_options()["seg_supported"]
    if seg_max > sup:
        pytest.skip("K=%d fits %d wavefronts per document" % (K, sup))
    ctx.set_infer_options(seg_tokens, seg_max)
    uw, uoff = _docs(V, seed=K)
    print("K", K, "seg_tokens", seg_tokens, "seg_max", seg_max, "docs", len(uoff) - 1, flush=True)
    th_b, ll_b, z_b = ctx.infer(uw, uoff, 12, infer_seed=3, mode=0, want_z=True)
    th_o, ll_o, z_o = m.infer(uw, uoff, 12, infer_seed=3, mode=0, seg_tokens=seg_tokens, seg_max=seg_max)
    assert np.array_equal(z_b, z_o), "z differs at %d tokens" % int((z_b != z_o).sum())
    assert np.array_equal(th_b, th_o)
    np.testing.assert_allclose(ll_b, ll_o, rtol=1e-9)
    th_p, ll_p, z_p = ctx.infer(uw, uoff, 12, infer_seed=3, mode=1, want_z=True)
    assert np.array_equal(z_p, z_b) and np.array_equal(th_p, th_b) and np.array_equal(ll_p, ll_b)


def test_segmented_document_alone_equals_in_batch():
    ctx, _, V = _setup(200, seed=9)
    ctx.set_infer_options(50, 8)
    uw, uoff = _docs(V, seed=4, n=40)
    th, ll, _ = ctx.infer(uw, uoff, 10)
    lens = np.diff(uoff.astype(np.int64))
    for d in (int(np.argmax(lens)), 0, len(lens) - 1):
        b, e = int(uoff[d]), int(uoff[d + 1])
        th1, ll1, _ = ctx.infer(uw[b:e], np.array([0, e - b], np.uint64), 10)
        assert np.array_equal(th1[0], th[d]) and ll1[0] == ll[d]


def test_segmented_serving_equals_infer():
    """The server takes the context's segmentation; split into steps and mixed
    with other requests, every result equals lda_infer's."""
    ctx, _, V = _setup(200, seed=10)
    ctx.set_infer_options(64, 8)
    reqs = []
    rng = np.random.default_rng(2)
    for r in range(10):
        uw, uoff = _docs(V, seed=100 + r, n=int(rng.integers(1, 8)))
        reqs.append((uw, uoff))
    with _native.Server(ctx, hbm_budget_bytes=32 << 20, iterations=15, step_iters=4, step_tokens=300,
                        infer_seed=5) as srv:
        ids = [srv.submit(w, o) for w, o in reqs]
        res = [srv.wait(i, timeout_ms=60000) for i in ids]
        assert srv.stats()["steps"] > 1
    for (w, o), (th, ll, _) in zip(reqs, res):
        th0, ll0, _ = ctx.infer(w, o, 15, infer_seed=5)
        assert np.array_equal(th, th0) and np.array_equal(ll, ll0)


def test_segment_options_validation():
    ctx, _, _ = _setup(20, seed=1, n_docs=20, iters=1)
    with pytest.raises(ValueError):
        ctx.set_infer_options(64, 3)
    opts = ctx.infer_options()
    assert opts["seg_tokens"] == 0 and opts["seg_max"] == 1 and opts["seg_supported"] in (1, 2, 4, 8)
