"""Integer counts with an alpha >= 2^20 or a document over 2^20 tokens (which
tomotopy accepts): the K <= 2048 samplers keep float(n_dk) + alpha and
recover n_dk from it, exact only below 2^21, so these models sample on the
exact-count kernels (n_dk kept as integers).  Bit-exact against the oracle
under every schedule and in fold-in, at every topics-per-lane count."""
import numpy as np
import pytest

from oracle import oracle as O
import tomotopy_amd as tp
from tomotopy_amd import _native, corpus, synth

pytestmark = pytest.mark.gpu


def _corpus(n_docs):
    words, off, nw, _ = synth.workload("plain_k20", n_docs=n_docs)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw)
    return ws, offs, V


def _compare(ctx, m):
    zg, ndkg, nkwg, nkg = ctx.download_state()
    zo, ndko, nkwo, nko = m.state()
    assert np.array_equal(zg, zo), "z differs at %d tokens" % int((zg != zo).sum())
    assert np.array_equal(ndkg, ndko)
    assert np.array_equal(nkwg, nkwo)
    assert np.array_equal(nkg, nko)


@pytest.mark.parametrize("K", [2, 20, 100, 200, 500, 1000])
def test_large_alpha_bit_exact(K):
    """alpha = 2^21 at every topics-per-lane count (1 ... 16): none, partition
    and copy_merge identical to the oracle, LL to 1e-9, fold-in bit-exact."""
    ws, offs, V = _corpus(40)
    rng = np.random.default_rng(K)
    alpha = (rng.random(K) * 0.5 + 0.01).astype(np.float32)
    alpha[K // 2] = np.float32(2.0**21)
    ctx = _native.Context(K, V, alpha, 0.01, seed=K)
    ctx.upload_corpus(ws, offs)
    ctx.set_optim(0, 0)
    m = O.Model(K, V, alpha, 0.01, K, scan_order=1)
    m.set_optim(0, 0)
    m.set_corpus(ws, offs)
    _compare(ctx, m)
    for workers, scheme in ((1, 1), (3, 3), (2, 2)):
        print("K", K, "workers", workers, "scheme", scheme, flush=True)
        ctx.train(1, workers=workers, scheme=scheme)
        m.train(1, workers=workers, scheme=scheme)
        _compare(ctx, m)
        assert not ctx.check_invariants().any()
    # the document part sums lgamma(n_dk + 2^21) - lgamma(2^21) terms of about
    # 2.9e7 each that cancel to ~1e-3, so it is held to an absolute bound
    # (double rounding of 40 documents' terms), the rest to 1e-9 relative
    lg, lo = ctx.loglik(), m.loglik()
    assert abs(lg[0] - lo[0]) < 1e-5
    np.testing.assert_allclose(lg[1], lo[1], rtol=1e-9)
    e = int(offs[5])
    uw, uoff = ws[:e].copy(), offs[:6].astype(np.uint64)
    tho, llo, zo = m.infer(uw, uoff, 6, infer_seed=3)
    for mode in (0, 1):
        thg, llg, zg = ctx.infer(uw, uoff, 6, infer_seed=3, mode=mode, want_z=True)
        assert np.array_equal(zg, zo) and np.array_equal(thg, tho)
        np.testing.assert_allclose(llg, llo, rtol=1e-9, atol=1e-6)
    # the device-wide schedule on the exact-count kernels keeps the counts exact
    ctx.train(2, workers=0, scheme=0)
    assert not ctx.check_invariants().any()


def test_long_document_bit_exact():
    """A corpus holding one document of 2^20 + 3,000 tokens: the none schedule
    identical to the oracle, exact invariants after a device-wide sweep, and
    the fold-in of a document of that length bit-exact."""
    K = 20
    ws, offs, V = _corpus(30)
    rng = np.random.default_rng(7)
    n_long = (1 << 20) + 3000
    long_doc = np.sort(rng.integers(0, V, n_long).astype(np.uint32))
    ws2 = np.concatenate([ws, long_doc])
    offs2 = np.concatenate([offs, [offs[-1] + n_long]]).astype(np.uint64)
    ctx = _native.Context(K, V, 0.1, 0.01, seed=13)
    ctx.upload_corpus(ws2, offs2)
    ctx.set_optim(0, 0)
    m = O.Model(K, V, 0.1, 0.01, 13, scan_order=1)
    m.set_optim(0, 0)
    m.set_corpus(ws2, offs2)
    _compare(ctx, m)
    print("long document sweep", flush=True)
    ctx.train(1, workers=1, scheme=1)
    m.train(1, workers=1, scheme=1)
    _compare(ctx, m)
    np.testing.assert_allclose(ctx.loglik(), m.loglik(), rtol=1e-9)
    print("fold-in of the long document", flush=True)
    uw = np.concatenate([ws[: int(offs[2])], long_doc])
    uoff = np.array([0, offs[1], offs[2], offs[2] + n_long], np.uint64)
    tho, llo, zo = m.infer(uw, uoff, 2, infer_seed=5)
    thg, llg, zg = ctx.infer(uw, uoff, 2, infer_seed=5, mode=0, want_z=True)
    assert np.array_equal(zg, zo) and np.array_equal(thg, tho)
    np.testing.assert_allclose(llg, llo, rtol=1e-9)
    ctx.optimize_alpha()
    m.optimize_alpha()
    np.testing.assert_allclose(ctx.get_alpha(), m.alpha(), rtol=2e-6)
    ctx.train(1, workers=0, scheme=0)
    assert not ctx.check_invariants().any()


def test_large_alpha_model_surface_and_server():
    """LDAModel with alpha >= 2^20 trains, infers and serves; the server's
    results equal the batched fold-in."""
    words, off, nw, _ = synth.workload("plain_k20", n_docs=60)
    mdl = tp.LDAModel(k=10, alpha=float(2**20), eta=0.01, seed=2)
    for d in range(len(off) - 1):
        mdl.add_doc(["w%d" % i for i in words[off[d]:off[d + 1]]])
    mdl.train(3, workers=1)
    assert np.isfinite(mdl.ll_per_word)
    doc = mdl.make_doc(["w%d" % i for i in words[off[0]:off[1]]])
    th, ll = mdl.infer(doc, iterations=4)
    assert abs(float(np.sum(th)) - 1.0) < 1e-4
    ws, offs, V = _corpus(40)
    ctx = _native.Context(20, V, float(2**21), 0.01, seed=1)
    ctx.upload_corpus(ws, offs)
    ctx.train(2, workers=1, scheme=1)
    uw, uoff = ws[: int(offs[6])].copy(), offs[:7].astype(np.uint64)
    with _native.Server(ctx, hbm_budget_bytes=64 << 20, iterations=5, infer_seed=4) as srv:
        th_s, ll_s, _ = srv.wait(srv.submit(uw, uoff), timeout_ms=60000)
    th0, ll0, _ = ctx.infer(uw, uoff, 5, infer_seed=4, mode=0)
    assert np.array_equal(th_s, th0) and np.array_equal(ll_s, ll0)
