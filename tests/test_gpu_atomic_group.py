"""Grouped n_kw atomics of the device-wide sweeps (lda_set_atomic_group): a
wavefront sends the n_kw changes of a group of its tokens at the group's end.
It never rereads a column with changes pending (a document's words come in
one run each, and its last group is sent before the next document), so one
wavefront alone samples the same chain whatever the group: integer counts
bit for bit at K <= 2048 and above.  With the full grid the count invariants
stay exact for every group."""
import numpy as np
import pytest

from tomotopy_amd import _native, corpus, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def zipf_corpus():
    words, off, nw, _ = synth.workload("large_k1000", n_docs=1500)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw)
    return ws, offs, V


def _single_wave_state(ws, offs, V, K, group, sweeps=2):
    ctx = _native.Context(K, V, 0.05, 0.01, seed=41)
    ctx.upload_corpus(ws, offs)
    ctx.set_optim(0, 0)
    ctx.set_device_schedule(1, 1)
    ctx.set_atomic_group(group)
    assert ctx.resolve_schedule(0, 0) == (0, 1)
    ctx.train(sweeps, workers=0)
    assert not ctx.check_invariants().any()
    st = ctx.download_state()
    ctx.close()
    return st


@pytest.mark.parametrize("K", [200, 1000, 3000])
def test_single_wavefront_same_chain_for_every_group(zipf_corpus, K):
    ws, offs, V = zipf_corpus
    ref = _single_wave_state(ws, offs, V, K, 1)
    for group in (8, 64, 0):
        print("K", K, "group", group, flush=True)
        st = _single_wave_state(ws, offs, V, K, group)
        for a, b, name in zip(ref, st, ("z", "ndk", "nkw", "nk")):
            assert np.array_equal(a, b), (group, name)


@pytest.mark.parametrize("group", [1, 4, 64])
def test_full_grid_invariants(zipf_corpus, group):
    ws, offs, V = zipf_corpus
    ctx = _native.Context(500, V, 0.05, 0.01, seed=43)
    ctx.upload_corpus(ws, offs)
    ctx.set_device_schedule(0, 8)  # every resident workgroup
    ctx.set_atomic_group(group)
    for it in range(2):
        ctx.train(2, workers=0)
        flags = ctx.check_invariants()
        print("group", group, "call", it, "invariant flags", flags.tolist(), flush=True)
        assert not flags.any()
    assert np.isfinite(sum(ctx.loglik()))
    ctx.close()


def test_weighted_full_grid_and_bad_groups(zipf_corpus):
    ws, offs, V = zipf_corpus
    wts = corpus.term_weights(ws, offs, V, "pmi")
    ctx = _native.Context(200, V, 0.05, 0.01, seed=44, tw=2)
    ctx.upload_corpus(ws, offs, weights=wts)
    ctx.set_device_schedule(0, 8)
    ctx.set_atomic_group(64)
    ctx.train(3, workers=0)
    _, _, nkw, nk = ctx.download_state_f()
    assert np.all(nkw >= 0) and np.all(np.isfinite(nk))
    # n_k is rebuilt from n_kw after every weighted sweep
    np.testing.assert_allclose(nk, nkw.sum(axis=0), rtol=1e-4, atol=1e-2)
    for bad in (-1, 3, 128):
        with pytest.raises(ValueError):
            ctx.set_atomic_group(bad)
    ctx.close()
