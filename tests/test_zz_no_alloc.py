"""lda_infer_noalloc: the batched fold-in staged in the library's fixed device
scratch gives lda_infer(mode 0)'s results bit for bit, on two calls in a row
(the second smaller, so stale contents of the first are in the scratch), with
segmented documents in several buckets, and with word priors; arenas beyond
the scratch's 64 MiB are refused before any device work."""
import numpy as np
import pytest

from tomotopy_amd import _native, corpus, synth

pytestmark = pytest.mark.gpu


def _model(K, seed, priors=False):
    words, off, nw, _ = synth.workload("plain_k20", n_docs=150)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw)
    ctx = _native.Context(K, V, 0.1, 0.01, seed=seed)
    if priors:
        rng = np.random.default_rng(seed)
        ids = np.arange(0, V, 7, dtype=np.uint32)
        ctx.set_word_priors(ids, (rng.random((len(ids), K)) * 0.05).astype(np.float32))
    ctx.upload_corpus(ws, offs)
    ctx.train(10, workers=0)
    return ctx, V


def _docs(V, n, seed, hi):
    uw, uoff = synth.unseen_docs(seed, n, V, 20, np.log(80.0), 1.0, 3, hi, zipf_s=None)
    return corpus.csr_from_arrays(uw, uoff, np.arange(V))


def _same(ctx, uw, uoff, iters, seed):
    th0, ll0, z0 = ctx.infer(uw, uoff, iters, infer_seed=seed, mode=0, want_z=True)
    th1, ll1, z1 = ctx.infer_noalloc(uw, uoff, iters, infer_seed=seed, want_z=True)
    assert np.array_equal(th0, th1)
    assert np.array_equal(ll0, ll1)
    assert np.array_equal(z0, z1)


@pytest.mark.parametrize("K,priors,seg", [(50, False, 0), (200, False, 64), (20, True, 32)])
def test_noalloc_equals_batched_two_calls(K, priors, seg):
    ctx, V = _model(K, 3, priors)
    if seg:
        ctx.set_infer_options(seg, ctx.infer_options()["seg_supported"])
    big = _docs(V, 60, 11, 2000)
    small = _docs(V, 7, 12, 300)
    print("K", K, "priors", priors, "seg", seg, "tokens", len(big[0]), len(small[0]), flush=True)
    _same(ctx, *big, 15, 5)
    _same(ctx, *small, 15, 6)
    _same(ctx, *big, 4, 7)


def test_noalloc_refuses_oversize_arena():
    ctx, V = _model(2048, 4)
    n = 9000  # theta alone: 9000 x 2048 x 4 B = 70 MiB
    uw = np.zeros(n, dtype=np.uint32)
    uoff = np.arange(n + 1, dtype=np.uint64)
    with pytest.raises(ValueError, match="scratch"):
        ctx.infer_noalloc(uw, uoff, 5)
