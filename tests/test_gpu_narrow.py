"""Narrow (u16) n_kw columns in the device-wide sweep.  With one worker of one
wavefront the device-wide schedule is sequential, so narrow and int32
columns must give the same topics and counts bit for bit; with the full
grid the exact count invariants must hold after every call; counts that do
not fit 15 bits make the call run on int32 columns."""
import numpy as np
import pytest

from tomotopy_amd import _native, corpus, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def zipf_corpus():
    words, off, nw, _ = synth.workload("large_k1000", n_docs=3000)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw)
    cf = np.bincount(ws[ws < V], minlength=V)
    print("tokens", len(ws), "V", V, "words with >= 2^15 tokens", int((cf >= 1 << 15).sum()), flush=True)
    assert (cf >= 1 << 15).any() and (cf < 1 << 15).any()
    return ws, offs, V


def _state(ctx):
    return ctx.download_state()


@pytest.mark.parametrize("K", [200, 1000])
def test_narrow_equals_int32_single_worker(zipf_corpus, K):
    ws, offs, V = zipf_corpus
    out = []
    for narrow in (True, False):
        ctx = _native.Context(K, V, 0.05, 0.01, seed=31)
        ctx.upload_corpus(ws, offs)
        ctx.set_optim(0, 0)
        ctx.set_device_schedule(1, 1)
        ctx.set_narrow_columns(narrow)
        assert ctx.resolve_schedule(0, 0) == (0, 1)
        ctx.train(2, workers=0)
        assert ctx.kernel_stats()["narrow"] == narrow
        print("K", K, "narrow", narrow, "kernel ms", ctx.kernel_stats()["ms"], flush=True)
        out.append(_state(ctx))
        assert not ctx.check_invariants().any()
        ctx.close()
    for a, b, name in zip(out[0], out[1], ("z", "ndk", "nkw", "nk")):
        assert np.array_equal(a, b), name


def test_narrow_full_grid_invariants(zipf_corpus):
    ws, offs, V = zipf_corpus
    ctx = _native.Context(1000, V, 0.05, 0.01, seed=32)
    ctx.upload_corpus(ws, offs)
    ctx.set_device_schedule(0, 8)  # every resident workgroup: most concurrent updates of the u16 cells
    for it in range(3):
        ctx.train(2, workers=0)
        assert ctx.kernel_stats()["narrow"]
        flags = ctx.check_invariants()
        print("call", it, "invariant flags", flags.tolist(), flush=True)
        assert not flags.any()


def test_counts_beyond_15_bits_run_on_int32(zipf_corpus):
    ws, offs, V = zipf_corpus
    K = 20
    ctx = _native.Context(K, V, 0.1, 0.01, seed=33)
    ctx.upload_corpus(ws, offs)
    _, _, nkw, nk = ctx.download_state(z=False, ndk=False)
    nkw = nkw.copy()
    nkw[V - 1, 0] += 40000  # a rare word's cell far above its token count
    nk = nk.copy()
    nk[0] += 40000
    ctx.set_counts(nkw, nk)
    ctx.train(1, workers=0)
    assert not ctx.kernel_stats()["narrow"]
    _, _, nkw2, _ = ctx.download_state(z=False, ndk=False)
    assert nkw2[V - 1, 0] >= 40000
