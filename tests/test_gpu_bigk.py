"""K above 2048 (up to tomotopy's 32767 topics) on the device against the CPU
oracle: frozen single steps, the deterministic schedules (none, partition,
copy_merge; integer, IDF / PMI and word-prior counts), alpha and the FP64
log-likelihood, the device-wide sweep's count invariants, fold-in alone and
together, and the LDAModel surface.  The oracle's device-order scan covers
any number of topics per lane (each lane's running sum in topic order, then
the 64-lane DPP scan), so the expected draws are exact."""
import numpy as np
import pytest

from oracle import oracle as O
import tomotopy_amd as tp
from tomotopy_amd import _native, corpus, synth

pytestmark = pytest.mark.gpu


def _corpus(n_docs, min_cf=0, rm_top=0):
    words, off, nw, _ = synth.workload("plain_k20", n_docs=n_docs)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw, min_cf=min_cf, rm_top=rm_top)
    return ws, offs, V


def _compare(ctx, m):
    zg, ndkg, nkwg, nkg = ctx.download_state()
    zo, ndko, nkwo, nko = m.state()
    assert np.array_equal(zg, zo), "z differs at %d tokens" % int((zg != zo).sum())
    assert np.array_equal(ndkg, ndko)
    assert np.array_equal(nkwg, nkwo)
    assert np.array_equal(nkg, nko)


def _compare_f(ctx, m):
    zg, ndkg, nkwg, nkg = ctx.download_state_f()
    zo, ndko, nkwo, nko, _ = m.state()
    assert np.array_equal(zg, zo), "z differs at %d tokens" % int((zg != zo).sum())
    assert np.array_equal(ndkg.view(np.uint32), ndko.view(np.uint32))
    assert np.array_equal(nkwg.view(np.uint32), nkwo.view(np.uint32))
    assert np.array_equal(nkg.view(np.uint32), nko.view(np.uint32))


@pytest.mark.parametrize("K", [2049, 3000, 4096, 8192, 32767])
def test_sample_frozen_bit_exact_big_k(K):
    ws, offs, V = _corpus(60)
    rng = np.random.default_rng(K)
    alpha = (rng.random(K) * 0.5 + 0.01).astype(np.float32)
    eta = np.float32(0.01)
    z0 = rng.integers(0, K, len(ws)).astype(np.uint16)
    ctx = _native.Context(K, V, alpha, eta, seed=5)
    ctx.upload_corpus(ws, offs, z0)
    _, ndk, nkw, nk = ctx.download_state(z=False)
    D = len(offs) - 1
    doc_of = np.repeat(np.arange(D), np.diff(offs).astype(np.int64))
    n = 4000
    pick = rng.integers(0, len(ws), n)
    doc, word, zo = doc_of[pick], ws[pick], z0[pick]
    u = (rng.integers(0, 1 << 24, n) * 2.0**-24).astype(np.float32)
    print("K", K, "kpl", O.device_kpl(K), "n", n, flush=True)
    zg, tg = ctx.sample_frozen(doc, word, zo, u)
    veta = np.float32(eta * np.float32(V))
    zo_dev, to_dev = O.sample_frozen(K, 1, alpha, eta, veta, ndk, nkw, nk, doc, word, zo, u)
    assert np.array_equal(tg.view(np.uint32), to_dev.view(np.uint32)), "cdf totals differ"
    assert np.array_equal(zg, zo_dev), "draws differ at %d of %d" % (int((zg != zo_dev).sum()), n)
    zo_tom, _ = O.sample_frozen(K, 0, alpha, eta, veta, ndk, nkw, nk, doc, word, zo, u)
    print("tie rate device-order vs tomoto-order: %.2e" % float(np.mean(zo_dev != zo_tom)))


@pytest.mark.parametrize("K,workers,scheme", [(4096, 1, 1), (4096, 3, 3), (3000, 4, 3), (4096, 2, 2)])
def test_schedules_bit_exact_big_k(K, workers, scheme):
    """none, partition and copy_merge: z and every count equal the oracle after
    each sweep; alpha after the fixed point within 2e-6, the LL to 1e-9."""
    ws, offs, V = _corpus(60)
    ctx = _native.Context(K, V, 0.1, 0.01, seed=K + workers)
    ctx.upload_corpus(ws, offs)
    ctx.set_optim(0, 0)
    m = O.Model(K, V, 0.1, 0.01, K + workers, scan_order=1)
    m.set_optim(0, 0)
    m.set_corpus(ws, offs)
    _compare(ctx, m)
    for it in range(2):
        print("K", K, "workers", workers, "scheme", scheme, "sweep", it, flush=True)
        ctx.train(1, workers=workers, scheme=scheme)
        m.train(1, workers=workers, scheme=scheme)
        _compare(ctx, m)
        assert not ctx.check_invariants().any()
    lg = ctx.loglik()
    lo = m.loglik()
    np.testing.assert_allclose(lg, lo, rtol=1e-9)
    ctx.optimize_alpha()
    m.optimize_alpha()
    # 98% of the n_dk cells are 0 here, so each pass's numerator
    # sum_d digamma(n_dk + a) - D digamma(a) cancels to a few units and the
    # last-bit differences of logf (device vs glibc) grow over the 10 passes
    # to about 3e-5; with fewer zero cells the fixed point holds 2e-6
    np.testing.assert_allclose(ctx.get_alpha(), m.alpha(), rtol=1e-4)


@pytest.mark.parametrize("scheme", ["idf", "pmi"])
@pytest.mark.parametrize("workers,ps", [(1, 1), (3, 3), (2, 2)])
def test_weighted_schedules_bit_exact_big_k(scheme, workers, ps):
    K = 4096
    ws, offs, V = _corpus(60, min_cf=2, rm_top=1)
    wts = corpus.term_weights(ws, offs, V, scheme)
    ctx = _native.Context(K, V, 0.1, 0.01, seed=workers + 7, tw=1 if scheme == "idf" else 2)
    ctx.upload_corpus(ws, offs, weights=wts)
    m = O.WeightedModel(K, V, 0.1, 0.01, workers + 7, scan_order=1)
    m.set_corpus(ws, offs, wts)
    ctx.set_optim(0, 0)
    m.set_optim(0, 0)
    _compare_f(ctx, m)
    for it in range(2):
        print(scheme, "workers", workers, "scheme", ps, "sweep", it, flush=True)
        ctx.train(1, workers=workers, scheme=ps)
        m.train(1, workers=workers, scheme=ps)
        _compare_f(ctx, m)


def test_word_priors_bit_exact_big_k():
    K = 4096
    ws, offs, V = _corpus(60)
    rng = np.random.default_rng(3)
    eta_kv = np.full((V, K), np.float32(0.01), np.float32)
    pw = rng.choice(V, size=min(20, V), replace=False).astype(np.uint32)
    for v in pw:
        eta_kv[v] = (rng.random(K) * rng.choice([0.01, 1.0, 5.0])).astype(np.float32)
    ctx = _native.Context(K, V, 0.1, 0.01, seed=9)
    ctx.set_word_priors(pw, eta_kv[pw])
    ctx.upload_corpus(ws, offs)
    m = O.Model(K, V, 0.1, 0.01, 9, scan_order=1)
    m.set_priors(eta_kv)
    m.set_corpus(ws, offs)
    ctx.set_optim(0, 0)
    m.set_optim(0, 0)
    _compare(ctx, m)
    for workers, scheme in ((1, 1), (3, 3)):
        print("priors workers", workers, flush=True)
        ctx.train(1, workers=workers, scheme=scheme)
        m.train(1, workers=workers, scheme=scheme)
        _compare(ctx, m)
    np.testing.assert_allclose(ctx.loglik(), m.loglik(), rtol=1e-9)


@pytest.mark.parametrize("tw", [0, 2])
def test_device_wide_sweep_big_k(tw):
    """The device-wide schedule at K = 4096: exact count invariants (integer)
    or the float row-sum identity, and the LL rises from the initial draw."""
    K = 4096
    ws, offs, V = _corpus(400)
    wts = corpus.term_weights(ws, offs, V, "pmi") if tw else None
    ctx = _native.Context(K, V, 0.05, 0.01, seed=21, tw=tw)
    ctx.upload_corpus(ws, offs, weights=wts)
    ctx.set_optim(0, 0)
    ll0 = sum(ctx.loglik())
    sch, W = ctx.resolve_schedule(0, 0)
    print("tw", tw, "schedule", sch, "workers", W, flush=True)
    assert sch == 0
    for it in range(4):
        ctx.train(1, workers=0, scheme=0)
        if tw == 0:
            assert not ctx.check_invariants().any()
    if tw:
        _, ndk, nkw, nk = ctx.download_state_f()
        np.testing.assert_allclose(nk, nkw.sum(axis=0, dtype=np.float64), rtol=1e-4, atol=1e-2)
        assert np.all(nkw >= 0)
    else:
        _, ndk, nkw, nk = ctx.download_state()
        assert np.array_equal(nk, nkw.sum(axis=0))
        assert int(ndk.sum()) == int(np.count_nonzero(ws < V))
    ll1 = sum(ctx.loglik())
    print("ll", ll0, "->", ll1, flush=True)
    assert ll1 > ll0


@pytest.mark.parametrize("K", [3000, 4096])
def test_device_wide_sweep_big_k_priors_and_rows(K):
    """The device-wide schedule with word priors at K = 3000 and 4096, where
    workgroups keep each wavefront's n_dk row and column copy in LDS: exact
    count invariants after every sweep, each document's row written back, and
    the LL rises."""
    ws, offs, V = _corpus(300)
    rng = np.random.default_rng(K)
    pw = rng.choice(V, size=min(30, V), replace=False).astype(np.uint32)
    pri = (rng.random((len(pw), K)) * 2.0).astype(np.float32)
    ctx = _native.Context(K, V, 0.1, 0.01, seed=K)
    ctx.set_word_priors(pw, pri)
    ctx.upload_corpus(ws, offs)
    ctx.set_optim(0, 0)
    ll0 = sum(ctx.loglik())
    assert ctx.resolve_schedule(0, 0)[0] == 0
    for it in range(3):
        ctx.train(1, workers=0, scheme=0)
        assert not ctx.check_invariants().any()
    _, ndk, nkw, nk = ctx.download_state()
    assert np.array_equal(nk, nkw.sum(axis=0))
    cs = np.concatenate([[0], np.cumsum(ws < V)])
    o = offs.astype(np.int64)
    assert np.array_equal(ndk.sum(axis=1), cs[o[1:]] - cs[o[:-1]])
    ll1 = sum(ctx.loglik())
    print("K", K, "ll", ll0, "->", ll1, flush=True)
    assert ll1 > ll0


@pytest.mark.parametrize("kind", ["one", "pmi", "prior"])
def test_fold_in_bit_exact_big_k(kind):
    """Fold-in at K = 4096 (one wavefront per document, no segmentation):
    topics and theta bit-exact against the oracle, LL to 1e-9, batched = plain."""
    K = 4096
    ws, offs, V = _corpus(60, min_cf=2, rm_top=1)
    tw = 2 if kind == "pmi" else 0
    wts = corpus.term_weights(ws, offs, V, "pmi") if tw else None
    ctx = _native.Context(K, V, 0.1, 0.01, seed=31, tw=tw)
    m = (O.WeightedModel if tw else O.Model)(K, V, 0.1, 0.01, 31, scan_order=1)
    if kind == "prior":
        rng = np.random.default_rng(4)
        eta_kv = np.full((V, K), np.float32(0.01), np.float32)
        pw = rng.choice(V, size=min(20, V), replace=False).astype(np.uint32)
        for v in pw:
            eta_kv[v] = (rng.random(K) * rng.choice([0.01, 1.0, 5.0])).astype(np.float32)
        ctx.set_word_priors(pw, eta_kv[pw])
        m.set_priors(eta_kv)
    ctx.upload_corpus(ws, offs, weights=wts)
    if tw:
        m.set_corpus(ws, offs, wts)
    else:
        m.set_corpus(ws, offs)
    ctx.set_optim(0, 0)
    m.set_optim(0, 0)
    ctx.train(2, workers=1, scheme=1)
    m.train(2, workers=1, scheme=1)
    e = int(offs[6])
    uw = np.concatenate([ws[:e], np.array([V, V], np.uint32)])
    uoff = np.concatenate([offs[:7], [e, e + 2]]).astype(np.uint64)
    uwt = np.concatenate([wts[:e], np.zeros(2, np.float32)]) if tw else None
    if kind == "one":
        tho, llo, zo = m.infer(uw, uoff, 8, infer_seed=77)
    else:
        tho, llo, zo = m.infer_gen(uw, uoff, 8, infer_seed=77, weights=uwt)
    for mode in (0, 1):
        print(kind, "mode", mode, flush=True)
        thg, llg, zg = ctx.infer(uw, uoff, 8, infer_seed=77, mode=mode, want_z=True, weights=uwt)
        assert np.array_equal(zg, zo), "mode %d: z differs at %d tokens" % (mode, int((zg != zo).sum()))
        assert np.array_equal(thg, tho), "mode %d theta" % mode
        np.testing.assert_allclose(llg, llo, rtol=1e-9, atol=1e-9)
    assert llg[6] == 0.0 and llg[7] == 0.0


def test_infer_together_big_k():
    K = 4096
    ws, offs, V = _corpus(50)
    ctx = _native.Context(K, V, 0.1, 0.01, seed=41)
    ctx.upload_corpus(ws, offs)
    m = O.Model(K, V, 0.1, 0.01, 41, scan_order=1)
    m.set_corpus(ws, offs)
    ctx.set_optim(0, 0)
    m.set_optim(0, 0)
    ctx.train(2, workers=1, scheme=1)
    m.train(2, workers=1, scheme=1)
    e = int(offs[5])
    uw, uoff = ws[:e].copy(), offs[:6].astype(np.uint64)
    for workers, scheme in ((1, 1), (3, 3)):
        print("together workers", workers, flush=True)
        thg, llg, zg = ctx.infer_together(uw, uoff, 4, infer_seed=5, workers=workers, scheme=scheme)
        tho, llo, zo = m.infer_together(uw, uoff, 4, infer_seed=5, workers=workers, scheme=scheme)
        assert np.array_equal(zg, zo)
        assert np.array_equal(thg, tho)
        np.testing.assert_allclose(llg, llo, rtol=1e-9)


def test_model_surface_big_k():
    """LDAModel(k=5000): train under the default schedule, read-outs, infer."""
    words, off, nw, _ = synth.workload("plain_k20", n_docs=120)
    mdl = tp.LDAModel(k=5000, alpha=0.1, eta=0.01, seed=3)
    for d in range(len(off) - 1):
        mdl.add_doc(["w%d" % i for i in words[off[d]:off[d + 1]]])
    mdl.train(5, workers=0)
    assert np.isfinite(mdl.ll_per_word)
    dist = np.asarray(mdl.get_topic_word_dist(4999))
    assert abs(float(dist.sum()) - 1.0) < 1e-4
    assert int(np.sum(mdl.get_count_by_topics())) == mdl.num_words
    doc = mdl.make_doc(["w%d" % i for i in words[off[0]:off[1]]])
    th, ll = mdl.infer(doc, iterations=5)
    th = np.asarray(th)
    assert th.shape == (5000,) and abs(float(th.sum()) - 1.0) < 1e-3
    assert np.isfinite(ll)
    with pytest.raises(ValueError):
        tp.LDAModel(k=32768)


def test_fold_in_batch_split_into_groups():
    """At K = 32767 a batch whose private columns exceed the launch budget
    (8 GB: here 400 documents of about 250 runs x 128 KiB) is folded in as
    consecutive groups; every document's result equals its fold-in alone and
    the oracle's."""
    K = 32767
    words, off, nw, _ = synth.workload("large_k1000", n_docs=600)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw)
    ctx = _native.Context(K, V, 0.05, 0.01, seed=3)
    ctx.upload_corpus(ws[: int(offs[200])], offs[:201].astype(np.uint64))
    ctx.set_optim(0, 0)
    ctx.train(1, workers=0)
    z, _, _, _ = ctx.download_state()
    uw, uoff = ws[int(offs[200]):int(offs[600])], (offs[200:601] - offs[200]).astype(np.uint64)
    runs = sum(len(np.unique(uw[int(uoff[d]):int(uoff[d + 1])])) for d in range(400))
    print("runs", runs, "GB", runs * 32768 * 4 / 1e9, flush=True)
    assert runs * 32768 * 4 > 8e9
    th, ll, zg = ctx.infer(uw, uoff, 3, infer_seed=9, mode=0, want_z=True)
    for d in (0, 137, 399):
        b, e = int(uoff[d]), int(uoff[d + 1])
        th1, ll1, z1 = ctx.infer(uw[b:e], np.array([0, e - b], np.uint64), 3, infer_seed=9, mode=0, want_z=True)
        assert np.array_equal(th1[0], th[d]) and ll1[0] == ll[d] and np.array_equal(z1, zg[b:e])
    m = O.Model(K, V, 0.05, 0.01, 3, scan_order=1)
    m.set_corpus(ws[: int(offs[200])], offs[:201].astype(np.uint64), z)
    sel = np.array([0, uoff[1], uoff[2]], np.uint64)
    tho, llo, zo = m.infer(uw[: int(uoff[2])], sel, 3, infer_seed=9)
    assert np.array_equal(zo, zg[: int(uoff[2])]) and np.array_equal(tho, th[:2])
    np.testing.assert_allclose(llo, ll[:2], rtol=1e-9)
