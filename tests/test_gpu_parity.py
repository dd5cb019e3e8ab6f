"""Device sampler vs the CPU oracle (bit-exact where the schedule is deterministic)."""
import numpy as np
import pytest

from oracle import oracle as O
from tomotopy_amd import _native, corpus, synth

pytestmark = pytest.mark.gpu


def _corpus(name, n_docs):
    words, off, nw, w = synth.workload(name, n_docs=n_docs)
    ws, offs, real_v, _ = corpus.prepare_arrays(words, off, nw)
    return ws, offs, real_v


def test_device_visible():
    print("devices:", _native.device_count(), flush=True)
    assert _native.device_count() >= 1


@pytest.mark.parametrize("K", [20, 100, 200, 500, 1000, 2048, 7, 65, 129])
def test_sample_frozen_bit_exact(K):
    """One getZLikelihoods + prefixSum + sampleFromDiscreteAcc step on a frozen state."""
    ws, offs, V = _corpus("plain_k20", 60)
    rng = np.random.default_rng(K)
    alpha = (rng.random(K) * 0.5 + 0.01).astype(np.float32)
    eta = np.float32(0.01)
    z0 = rng.integers(0, K, len(ws)).astype(np.uint16)
    ctx = _native.Context(K, V, alpha, eta, seed=5)
    ctx.upload_corpus(ws, offs, z0)
    _, ndk, nkw, nk = ctx.download_state(z=False)
    D = len(offs) - 1
    doc_of = np.repeat(np.arange(D), np.diff(offs).astype(np.int64))
    n = 20000
    pick = rng.integers(0, len(ws), n)
    doc, word, zo = doc_of[pick], ws[pick], z0[pick]
    u = (rng.integers(0, 1 << 24, n) * 2.0**-24).astype(np.float32)
    print("K", K, "n", n, flush=True)
    zg, tg = ctx.sample_frozen(doc, word, zo, u)
    veta = np.float32(eta * np.float32(V))
    zo_dev, to_dev = O.sample_frozen(K, 1, alpha, eta, veta, ndk, nkw, nk, doc, word, zo, u)
    zo_tom, _ = O.sample_frozen(K, 0, alpha, eta, veta, ndk, nkw, nk, doc, word, zo, u)
    assert np.array_equal(tg.view(np.uint32), to_dev.view(np.uint32)), "cdf totals differ"
    assert np.array_equal(zg, zo_dev)
    rate = float(np.mean(zo_dev != zo_tom))
    print("tie rate device-order vs tomoto-order: %.2e" % rate)
    # measured on the oracle's two scan orders over 200,000 draws: 0 for
    # K <= 100, 2.0e-5 (K=200), 2.5e-5 (500), 6.0e-5 (1000), 2.3e-4 (2048)
    assert rate < 1e-3


def _compare_state(ctx, m):
    zg, ndkg, nkwg, nkg = ctx.download_state()
    zo, ndko, nkwo, nko = m.state()
    assert np.array_equal(zg, zo), "z differs at %d tokens" % int((zg != zo).sum())
    assert np.array_equal(ndkg, ndko)
    assert np.array_equal(nkwg, nkwo)
    assert np.array_equal(nkg, nko)


@pytest.mark.parametrize("workers", [1, 3, 8])
def test_deterministic_schedule_bit_exact(workers):
    """none (workers=1) and partition (workers=W): z and all counts equal the oracle after every sweep."""
    ws, offs, V = _corpus("plain_k20", 80)
    K = 20
    ctx = _native.Context(K, V, 0.1, 0.01, seed=11)
    ctx.upload_corpus(ws, offs)
    ctx.set_optim(0, 0)
    m = O.Model(K, V, 0.1, 0.01, 11, scan_order=1)
    m.set_optim(0, 0)
    m.set_corpus(ws, offs)
    _compare_state(ctx, m)  # initial draw
    for it in range(3):
        print("sweep", it, "workers", workers, flush=True)
        ctx.train(1, workers=workers, scheme=3)
        m.train(1, workers=workers)
        _compare_state(ctx, m)
        assert not ctx.check_invariants().any()


@pytest.mark.parametrize("scheme", [3, 2])
def test_workers0_runs_the_explicit_scheme(scheme):
    """train(workers=0, parallel=PARTITION|COPY_MERGE) runs that scheme with the
    device's worker count (tomotopy: workers=0 means the hardware's count,
    TopicModel.hpp:686); bit-exact against the oracle at the same W."""
    ws, offs, V = _corpus("plain_k20", 300)
    K = 20
    ctx = _native.Context(K, V, 0.1, 0.01, seed=17)
    ctx.upload_corpus(ws, offs)
    ctx.set_optim(0, 0)
    sch, W = ctx.resolve_schedule(0, scheme)
    print("scheme", scheme, "resolved", sch, "W", W, flush=True)
    assert sch == scheme and W > 1
    assert ctx.resolve_schedule(0, 0)[0] == 0  # DEFAULT with workers=0: the device-wide schedule
    assert ctx.resolve_schedule(4, 0) == (3, 4)  # DEFAULT with workers: partition, as tomoto's LDA
    m = O.Model(K, V, 0.1, 0.01, 17, scan_order=1)
    m.set_optim(0, 0)
    m.set_corpus(ws, offs)
    for it in range(3):
        ctx.train(1, workers=0, scheme=scheme)
        m.train(1, workers=W, scheme=scheme)
        _compare_state(ctx, m)
        assert not ctx.check_invariants().any()


def test_partition_graph_replay_follows_changes():
    """The partition rounds replay a captured graph: changing W, alpha and the
    sweep number between calls gives the oracle's result every time."""
    ws, offs, V = _corpus("plain_k20", 90)
    K = 20
    ctx = _native.Context(K, V, 0.1, 0.01, seed=19)
    ctx.upload_corpus(ws, offs)
    ctx.set_optim(0, 0)
    m = O.Model(K, V, 0.1, 0.01, 19, scan_order=1)
    m.set_optim(0, 0)
    m.set_corpus(ws, offs)
    new_alpha = np.linspace(0.05, 0.3, K).astype(np.float32)
    for it, W in enumerate([3, 3, 5, 3, 5, 5]):
        if it == 3:
            ctx.set_alpha(new_alpha)
            m.set_alpha(new_alpha)
        print("sweep", it, "W", W, flush=True)
        ctx.train(1, workers=W, scheme=3)
        m.train(1, workers=W)
        _compare_state(ctx, m)


@pytest.mark.parametrize("workers", [2, 4])
def test_copy_merge_schedule_bit_exact(workers):
    """copy_merge (private n_kw copies, merged by summing deltas): z and all counts equal the oracle."""
    ws, offs, V = _corpus("plain_k20", 70)
    K = 20
    ctx = _native.Context(K, V, 0.1, 0.01, seed=13)
    ctx.upload_corpus(ws, offs)
    ctx.set_optim(0, 0)
    m = O.Model(K, V, 0.1, 0.01, 13, scan_order=1)
    m.set_optim(0, 0)
    m.set_corpus(ws, offs)
    for it in range(3):
        print("copy_merge sweep", it, "workers", workers, flush=True)
        ctx.train(1, workers=workers, scheme=2)
        m.train(1, workers=workers, scheme=2)
        _compare_state(ctx, m)
        assert not ctx.check_invariants().any()


@pytest.mark.parametrize("K", [20, 200])
def test_training_with_alpha_optimisation_matches_oracle(K):
    """alpha's fixed point runs on the device (count histograms, no host round
    trip); the oracle sums the same terms document by document."""
    ws, offs, V = _corpus("plain_k20", 60)
    ctx = _native.Context(K, V, 0.1, 0.01, seed=3)
    ctx.upload_corpus(ws, offs)
    ctx.set_optim(5, 0)
    m = O.Model(K, V, 0.1, 0.01, 3, scan_order=1)
    m.set_optim(5, 0)
    m.set_corpus(ws, offs)
    ctx.train(5, workers=1, scheme=1)
    m.train(5, workers=1)
    assert not np.array_equal(ctx.get_alpha(), np.full(K, np.float32(0.1)))  # alpha moved
    np.testing.assert_allclose(ctx.get_alpha(), m.alpha(), rtol=2e-6)
    _compare_state(ctx, m)


def test_loglik_fp64_matches_oracle():
    ws, offs, V = _corpus("plain_k20", 100)
    K = 20
    ctx = _native.Context(K, V, 0.1, 0.01, seed=9)
    ctx.upload_corpus(ws, offs)
    m = O.Model(K, V, 0.1, 0.01, 9, scan_order=1)
    m.set_corpus(ws, offs)
    ctx.train(4, workers=0)
    z, _, _, _ = ctx.download_state()
    m.set_corpus(ws, offs, z)
    m.set_alpha(ctx.get_alpha())
    gd, gr = ctx.loglik()
    od, orr = m.loglik(0)
    assert abs(gd - od) <= 1e-9 * abs(od)
    assert abs(gr - orr) <= 1e-9 * abs(orr)


@pytest.mark.parametrize("K", [20, 200, 1000])
def test_device_wide_sweep_invariants(K):
    ws, offs, V = _corpus("plain_k20", 300)
    ctx = _native.Context(K, V, 0.1, 0.01, seed=21)
    ctx.upload_corpus(ws, offs)
    for it in range(4):
        print("device-wide sweep", it, "K", K, flush=True)
        ctx.train(1, workers=0)
        f = ctx.check_invariants()
        assert not f.any(), f


def test_infer_batched_plain_and_oracle_agree():
    ws, offs, V = _corpus("plain_k20", 100)
    K = 20
    ctx = _native.Context(K, V, 0.1, 0.01, seed=4)
    ctx.upload_corpus(ws, offs)
    ctx.train(10, workers=0)
    z, _, _, _ = ctx.download_state()
    m = O.Model(K, V, ctx.get_alpha(), 0.01, 4, scan_order=1)
    m.set_corpus(ws, offs, z)
    uw, uoff = synth.unseen_docs(77, 30, V, 20, np.log(100.0), 1.0, 3, 600, zipf_s=None)
    uw_sorted, uoff = corpus.csr_from_arrays(uw, uoff, np.arange(V))
    th_b, ll_b, z_b = ctx.infer(uw_sorted, uoff, 20, infer_seed=0, mode=0, want_z=True)
    th_p, ll_p, z_p = ctx.infer(uw_sorted, uoff, 20, infer_seed=0, mode=1, want_z=True)
    th_o, ll_o, z_o = m.infer(uw_sorted, uoff, 20, infer_seed=0, mode=0)
    assert np.array_equal(z_b, z_o)
    assert np.array_equal(z_b, z_p)
    assert np.array_equal(th_b, th_o)
    np.testing.assert_allclose(ll_b, ll_o, rtol=1e-9)
    np.testing.assert_array_equal(ll_b, ll_p)


def test_infer_graph_replay_and_batch_independence():
    """The batched launch is a replayed hipGraph: replays with other batches give
    each document the same result as when it runs alone."""
    ws, offs, V = _corpus("plain_k20", 100)
    K = 20
    ctx = _native.Context(K, V, 0.1, 0.01, seed=8)
    ctx.upload_corpus(ws, offs)
    ctx.train(5, workers=0)
    uw, uoff = synth.unseen_docs(91, 40, V, 20, np.log(80.0), 1.0, 2, 900, zipf_s=None)
    uw, uoff = corpus.csr_from_arrays(uw, uoff, np.arange(V))
    th_all, ll_all, _ = ctx.infer(uw, uoff, 15)
    th_again, ll_again, _ = ctx.infer(uw, uoff, 15)
    assert np.array_equal(th_all, th_again) and np.array_equal(ll_all, ll_again)
    for d in (0, 7, 39):
        b, e = int(uoff[d]), int(uoff[d + 1])
        th1, ll1, _ = ctx.infer(uw[b:e], np.array([0, e - b], np.uint64), 15)
        assert np.array_equal(th1[0], th_all[d]) and ll1[0] == ll_all[d]


@pytest.mark.parametrize("exp_shift", [-20, -1, 0, 5, 10, 17, 24, 28])
def test_device_reciprocal_is_correctly_rounded(exp_shift):
    """The sampler's RN(1/b) (hardware reciprocal + one Newton step) equals
    IEEE division for all 2^23 mantissas of the binade."""
    assert _native.selftest_reciprocal(exp_shift) == 0
