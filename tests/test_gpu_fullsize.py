"""The device path at the benchmark workloads' full sizes, checked through
properties that hold at any size: the count invariants after sweeps of the
K=1000, 30M-token corpus (exact for integer counts; for PMI-weighted float
counts non-negativity, the row-sum identity and the document weight sums),
the FP64 log-likelihood rising from the initial state, and the K=200 fold-in
of 20,000 unseen documents (theta rows summing to one, batched equal to plain
and to one document alone, the segmented sampler's ll_per_word within 1% of
the sequential one)."""
import numpy as np
import pytest

from tomotopy_amd import _native, corpus, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def large_k1000():
    words, off, nw, _ = synth.workload("large_k1000")
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw)
    print("large_k1000:", len(offs) - 1, "docs", len(ws), "tokens, V", V, flush=True)
    return ws, offs, V


def _ll(ctx, n):
    d, r = ctx.loglik()
    return (d + r) / n


def test_large_k1000_device_wide_and_partition_invariants(large_k1000):
    ws, offs, V = large_k1000
    K = 1000
    n = int(np.count_nonzero(ws < V))
    ctx = _native.Context(K, V, 0.05, 0.01, seed=3)
    ctx.upload_corpus(ws, offs)
    ctx.set_optim(2, 0)  # the alpha fixed point runs after the second sweep
    assert not ctx.check_invariants().any()
    ll0 = _ll(ctx, n)
    for it in range(3):
        print("device-wide sweep", it, flush=True)
        ctx.train(1, workers=0)
        f = ctx.check_invariants()
        assert not f.any(), f
    print("partition sweep, 8 workers", flush=True)
    ctx.train(1, workers=8, scheme=3)
    assert not ctx.check_invariants().any()
    ll1 = _ll(ctx, n)
    alpha = ctx.get_alpha()
    print("ll_per_word %.5f -> %.5f" % (ll0, ll1), flush=True)
    assert np.isfinite(ll1) and ll1 > ll0
    assert np.all(alpha > 0) and not np.array_equal(alpha, np.full(K, np.float32(0.05)))


def test_large_k1000_pmi_invariants(large_k1000):
    ws, offs, V = large_k1000
    K = 1000
    wts = corpus.term_weights(ws, offs, V, "pmi")
    ctx = _native.Context(K, V, 0.05, 0.01, seed=4, tw=2)
    ctx.upload_corpus(ws, offs, weights=wts)
    wn = ctx.weighted_n()
    ll0 = _ll(ctx, wn)
    for it in range(2):
        print("PMI device-wide sweep", it, flush=True)
        ctx.train(1, workers=0)
    # float counts: non-negative, n_k the two-level float row sum of n_kw
    # (blocks of 256 words in word order, then the blocks in order), document
    # rows summing to the document weights up to float rounding
    z, ndk, nkw, nk = ctx.download_state_f()
    valid = ws < V
    assert z[valid].max() < K
    assert (ndk >= 0).all() and (nkw >= 0).all() and (nk >= 0).all()
    blocks = [np.cumsum(nkw[v0:v0 + 256], axis=0, dtype=np.float32)[-1] for v0 in range(0, V, 256)]
    seq = np.cumsum(np.array(blocks), axis=0, dtype=np.float32)[-1]
    assert np.array_equal(nk.view(np.uint32), seq.view(np.uint32))
    doc_of = np.repeat(np.arange(len(offs) - 1), np.diff(offs.astype(np.int64)))
    docw = np.bincount(doc_of[valid], weights=wts[valid].astype(np.float64), minlength=len(offs) - 1)
    np.testing.assert_allclose(ndk.sum(axis=1, dtype=np.float64), docw, rtol=2e-3, atol=1e-2)
    del ndk, nkw
    ll1 = _ll(ctx, wn)
    print("PMI ll per weight %.5f -> %.5f" % (ll0, ll1), flush=True)
    assert np.isfinite(ll1) and ll1 > ll0


def test_foldin_k200_full_batch():
    (tw, toff, nw), (uw, uoff) = synth.foldin_workload()
    ws, offs, V, order = corpus.prepare_arrays(tw, toff, nw)
    uws, uoffs = corpus.csr_from_arrays(uw, uoff, order)
    D = len(uoffs) - 1
    K, iters = 200, 20
    ctx = _native.Context(K, V, 0.1, 0.01, seed=77)
    ctx.upload_corpus(ws, offs)
    ctx.train(20, workers=0)
    ntok = int(np.count_nonzero(uws < V))
    print("foldin_k200:", D, "unseen docs", len(uws), "tokens", flush=True)
    ctx.set_infer_options(1024, 8)
    th, ll, _ = ctx.infer(uws, uoffs, iters, infer_seed=5)
    assert th.shape == (D, K) and np.all(np.isfinite(ll))
    np.testing.assert_allclose(th.sum(axis=1, dtype=np.float64), 1.0, rtol=1e-5)
    lens = np.diff(uoffs.astype(np.int64))
    pick = np.sort(np.random.default_rng(1).choice(D, 16, replace=False))
    pick = np.union1d(pick, [int(np.argmax(lens))])  # the longest document is split over 8 wavefronts
    sub_off = np.zeros(len(pick) + 1, np.uint64)
    sub_off[1:] = np.cumsum(lens[pick])
    sub_w = np.concatenate([uws[int(uoffs[d]):int(uoffs[d + 1])] for d in pick]).astype(np.uint32)
    th_p, ll_p, _ = ctx.infer(sub_w, sub_off, iters, infer_seed=5, mode=1)
    assert np.array_equal(th_p, th[pick]) and np.array_equal(ll_p, ll[pick])
    d = int(pick[0])
    th1, ll1, _ = ctx.infer(uws[int(uoffs[d]):int(uoffs[d + 1])], np.array([0, lens[d]], np.uint64), iters,
                            infer_seed=5)
    assert np.array_equal(th1[0], th[d]) and ll1[0] == ll[d]
    ctx.set_infer_options(0, 1)
    _, ll_seq, _ = ctx.infer(uws, uoffs, iters, infer_seed=5)
    a, b = float(ll.sum() / ntok), float(ll_seq.sum() / ntok)
    print("ll_per_word segmented %.5f sequential %.5f" % (a, b), flush=True)
    assert abs(a - b) <= 0.01 * abs(b)
