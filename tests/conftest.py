import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the C-ABI library on the GPU")
    config.addinivalue_line("markers", "slow: longer CPU-side check")


@pytest.fixture(scope="session")
def small_corpus():
    """plain_k20 generative corpus cut to 120 documents, prepared as tomoto does."""
    from tomotopy_amd import corpus, synth
    words, off, nw, w = synth.workload("plain_k20", n_docs=120)
    ws, offs, real_v, order = corpus.prepare_arrays(words, off, nw)
    return ws, offs, real_v


def gpu_ctx(K, V, alpha, eta, seed):
    from tomotopy_amd import _native
    return _native.Context(K, V, alpha, eta, seed)
