"""Edge cases of the device path against the oracle: the smallest and the
largest K (1, 2, 2048), a one-word vocabulary, and corpora with empty
documents and documents made only of removed words; training under the
deterministic schedules bit-exact with the count invariants, and fold-in
bit-exact."""
import numpy as np
import pytest

from oracle import oracle as O
from tomotopy_amd import _native, corpus, synth

pytestmark = pytest.mark.gpu


def _ragged_corpus(n_docs=60, rm_top=2):
    """A small corpus with empty documents and documents of removed words only
    (ids >= V), documents sorted by id as prepare() leaves them."""
    words, off, nw, _ = synth.workload("plain_k20", n_docs=n_docs)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw, min_cf=2, rm_top=rm_top)
    docs = [ws[int(offs[d]):int(offs[d + 1])] for d in range(len(offs) - 1)]
    docs.insert(3, np.zeros(0, np.uint32))
    docs.insert(10, np.array([V, V, V + 1], np.uint32))
    docs.append(np.zeros(0, np.uint32))
    out_off = np.zeros(len(docs) + 1, np.uint64)
    out_off[1:] = np.cumsum([len(d) for d in docs])
    return np.concatenate(docs).astype(np.uint32), out_off, V


def _same(ctx, m):
    zg, ndkg, nkwg, nkg = ctx.download_state()
    zo, ndko, nkwo, nko = m.state()
    assert np.array_equal(zg, zo), "z differs at %d tokens" % int((zg != zo).sum())
    assert np.array_equal(ndkg, ndko) and np.array_equal(nkwg, nkwo) and np.array_equal(nkg, nko)


@pytest.mark.parametrize("K", [1, 2, 2048])
@pytest.mark.parametrize("workers,scheme", [(1, 1), (3, 3), (2, 2)])
def test_extreme_k_schedules_bit_exact(K, workers, scheme):
    ws, offs, V = _ragged_corpus()
    ctx = _native.Context(K, V, 0.1, 0.01, seed=K + workers)
    ctx.upload_corpus(ws, offs)
    # alpha's fixed point with 2048 topics over 60 documents (almost every
    # n_dk is 0) cancels to a few units per topic and amplifies last-bit
    # differences of the two logf implementations inside digammaApprox; its
    # parity is held in the well-conditioned cases of test_gpu_parity
    interval = 2 if K < 2048 else 0
    ctx.set_optim(interval, 0)
    m = O.Model(K, V, 0.1, 0.01, K + workers, scan_order=1)
    m.set_optim(interval, 0)
    m.set_corpus(ws, offs)
    _same(ctx, m)
    for it in range(2):
        print("K", K, "workers", workers, "scheme", scheme, "sweep", it, flush=True)
        ctx.train(1, workers=workers, scheme=scheme)
        m.train(1, workers=workers, scheme=scheme)
        _same(ctx, m)
        assert not ctx.check_invariants().any()
    np.testing.assert_allclose(ctx.get_alpha(), m.alpha(), rtol=2e-6)
    if K == 1:
        z, _, _, _ = ctx.download_state()
        assert set(np.unique(z[ws < V]).tolist()) <= {0}


@pytest.mark.parametrize("K", [1, 2, 2048])
def test_extreme_k_device_wide_and_fold_in(K):
    ws, offs, V = _ragged_corpus()
    ctx = _native.Context(K, V, 0.1, 0.01, seed=5)
    ctx.upload_corpus(ws, offs)
    for it in range(3):
        print("device-wide K", K, "sweep", it, flush=True)
        ctx.train(1, workers=0)
        assert not ctx.check_invariants().any()
    z, _, _, _ = ctx.download_state()
    m = O.Model(K, V, ctx.get_alpha(), 0.01, 5, scan_order=1)
    m.set_corpus(ws, offs, z)
    e = int(offs[14])
    uw, uoff = ws[:e], offs[:15]
    tho, llo, zo = m.infer(uw, uoff, 6, infer_seed=3, mode=0)
    for mode in (0, 1):
        thg, llg, zg = ctx.infer(uw, uoff, 6, infer_seed=3, mode=mode, want_z=True)
        assert np.array_equal(zg, zo) and np.array_equal(thg, tho)
        np.testing.assert_allclose(llg, llo, rtol=1e-9, atol=1e-9)


def test_one_word_vocabulary():
    """V = 1: every valid token is the same word; schedules and fold-in still
    match the oracle."""
    rng = np.random.default_rng(1)
    lens = rng.integers(0, 12, 30)
    off = np.zeros(31, np.uint64)
    off[1:] = np.cumsum(lens)
    ws = np.zeros(int(off[-1]), np.uint32)
    K = 7
    ctx = _native.Context(K, 1, 0.1, 0.01, seed=2)
    ctx.upload_corpus(ws, off)
    m = O.Model(K, 1, 0.1, 0.01, 2, scan_order=1)
    m.set_corpus(ws, off)
    for x in (ctx, m):
        x.set_optim(0, 0)
    for it in range(3):
        ctx.train(1, workers=1, scheme=1)
        m.train(1, workers=1, scheme=1)
        _same(ctx, m)
    tho, llo, zo = m.infer(ws[:int(off[5])], off[:6], 5, infer_seed=1, mode=0)
    thg, llg, zg = ctx.infer(ws[:int(off[5])], off[:6], 5, infer_seed=1, mode=0, want_z=True)
    assert np.array_equal(zg, zo) and np.array_equal(thg, tho)
    np.testing.assert_allclose(llg, llo, rtol=1e-9, atol=1e-9)
