"""The README-scale workload at its full size (1,000 documents of mean
length 1,500, about 1.5M tokens, V=50,000 with a Zipf(1.07) background,
K=100, alpha 0.1, eta 0.01) under tomoto's partition schedule
(ParallelScheme.PARTITION, LDAModel.hpp:373-399, 436-489): z and every count
bit for bit equal to the CPU oracle's after each sweep at W = 8 and 64 and at
the device's own worker count, with the exact count invariants; then alpha
optimisation included, alpha within 2e-6 of the oracle's."""
import numpy as np
import pytest

from oracle import oracle as O
from tomotopy_amd import _native, corpus, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def readme_corpus():
    words, off, nw, _ = synth.workload("readme_k100")
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw)
    print("documents", len(offs) - 1, "tokens", len(ws), "V", V, flush=True)
    return ws, offs, V


def _compare_state(ctx, m):
    zg, ndkg, nkwg, nkg = ctx.download_state()
    zo, ndko, nkwo, nko = m.state()
    assert np.array_equal(zg, zo), "z differs at %d tokens" % int((zg != zo).sum())
    assert np.array_equal(ndkg, ndko)
    assert np.array_equal(nkwg, nkwo)
    assert np.array_equal(nkg, nko)


@pytest.mark.parametrize("workers", [8, 64, 0])
def test_readme_k100_partition_bit_exact(readme_corpus, workers):
    ws, offs, V = readme_corpus
    K = 100
    ctx = _native.Context(K, V, 0.1, 0.01, seed=1002)
    ctx.upload_corpus(ws, offs)
    ctx.set_optim(0, 0)
    sch, W = ctx.resolve_schedule(workers, 3)
    assert sch == 3
    m = O.Model(K, V, 0.1, 0.01, 1002, scan_order=1)
    m.set_optim(0, 0)
    m.set_corpus(ws, offs)
    for it in range(2):
        print("workers", workers, "W", W, "sweep", it, flush=True)
        ctx.train(1, workers=workers, scheme=3)
        m.train(1, workers=W, threads=8)
        _compare_state(ctx, m)
        assert not ctx.check_invariants().any()


def test_readme_k100_partition_with_alpha(readme_corpus):
    ws, offs, V = readme_corpus
    K = 100
    ctx = _native.Context(K, V, 0.1, 0.01, seed=7)
    ctx.upload_corpus(ws, offs)
    ctx.set_optim(2, 0)
    m = O.Model(K, V, 0.1, 0.01, 7, scan_order=1)
    m.set_optim(2, 0)
    m.set_corpus(ws, offs)
    ctx.train(2, workers=16, scheme=3)
    m.train(2, workers=16, threads=8)
    _compare_state(ctx, m)
    np.testing.assert_allclose(ctx.get_alpha(), m.alpha(), rtol=2e-6)
    assert not ctx.check_invariants().any()
