"""LDAModel.copy() and the "Train failed" error path on the GPU.

copy() must behave as tomoto's copy constructor (TopicModel.hpp:623-628):
the copy holds the same topics and counts bit for bit, including the float
counts of IDF / PMI models that training produced through clamped removals
and ordered merges, and the counts of an inference-only model (saved with
full=False, so it has no documents to rebuild them from).  A training error
surfaces as RuntimeError("Train failed", py_LDA.cpp:306-309), forced from the
host through a test hook that affects only the model it is set on."""
import os

import numpy as np
import pytest

import tomotopy_amd as tp

pytestmark = pytest.mark.gpu

SAMPLE = os.path.join(os.path.dirname(__file__), "golden", "sample.txt")


def _lines(n=None):
    with open(SAMPLE, encoding="utf-8") as f:
        out = [l.strip().split() for l in f]
    out = [ch for ch in out if ch]
    return out if n is None else out[:n]


def _model(lines, **kw):
    mdl = tp.LDAModel(**kw)
    for ch in lines:
        mdl.add_doc(ch)
    return mdl


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype.itemsize == 4 else a


def _state(mdl):
    mdl._cache = {}
    return {k: _bits(mdl._state(k)).copy() for k in ("z", "ndk", "nkw", "nk")}


def _same_state(a, b):
    for k in ("z", "ndk", "nkw", "nk"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("tw", [0, 1, 2])
def test_copy_is_exact_and_trains_alike(tw):
    lines = _lines()
    mdl = _model(lines[5:], tw=tw, min_df=2, rm_top=2, k=12, seed=11)
    mdl.train(30, workers=4, parallel=tp.ParallelScheme.PARTITION)
    cp = mdl.copy()
    print("tw", tw, "copied at step", mdl.global_step, flush=True)
    _same_state(_state(mdl), _state(cp))
    assert cp.ll_per_word == mdl.ll_per_word
    assert list(cp.alpha) == list(mdl.alpha) and cp.global_step == mdl.global_step
    # the same deterministic schedule continues identically on both
    mdl.train(5, workers=3, parallel=tp.ParallelScheme.PARTITION)
    cp.train(5, workers=3, parallel=tp.ParallelScheme.PARTITION)
    _same_state(_state(mdl), _state(cp))
    assert cp.ll_per_word == mdl.ll_per_word
    # and folds in unseen documents identically
    unseen = [mdl.make_doc(ch) for ch in lines[:5]]
    unseen_c = [cp.make_doc(ch) for ch in lines[:5]]
    d1, ll1 = mdl.infer(unseen, iterations=20)
    d2, ll2 = cp.infer(unseen_c, iterations=20)
    assert np.array_equal(np.asarray(d1, np.float32), np.asarray(d2, np.float32)) and ll1 == ll2


@pytest.mark.parametrize("tw", [0, 2])
def test_copy_of_inference_only_model(tw):
    lines = _lines()
    mdl = _model(lines[5:], tw=tw, min_df=2, rm_top=2, k=12, seed=5)
    mdl.train(30)
    light = tp.LDAModel.loads(mdl.saves(full=False))
    cp = light.copy()
    assert np.array_equal(_bits(cp._state("nkw")), _bits(mdl._state("nkw")))
    assert np.array_equal(_bits(cp._state("nk")), _bits(mdl._state("nk")))
    unseen = [mdl.make_doc(ch) for ch in lines[:5]]
    d0, ll0 = mdl.infer(unseen, iterations=20)
    d1, ll1 = light.infer([light.make_doc(ch) for ch in lines[:5]], iterations=20)
    d2, ll2 = cp.infer([cp.make_doc(ch) for ch in lines[:5]], iterations=20)
    print("tw", tw, "ll", ll0[:2], ll2[:2], flush=True)
    assert np.array_equal(np.asarray(d0, np.float32), np.asarray(d2, np.float32)) and ll0 == ll2
    assert np.array_equal(np.asarray(d1, np.float32), np.asarray(d2, np.float32)) and ll1 == ll2


def test_train_failed_is_runtime_error_and_local():
    lines = _lines()
    mdl = _model(lines[5:], tw=0, min_df=2, rm_top=2, k=8, seed=2)
    mdl.train(5)
    cp = mdl.copy()
    mdl._ctx.fail_at_step(7)
    with pytest.raises(RuntimeError, match="Train failed"):
        mdl.train(5)
    # together=True samples on a private copy of the model: the hook set on
    # the model does not reach it
    unseen = [mdl.make_doc(ch) for ch in lines[:5]]
    dists, ll = mdl.infer(unseen, iterations=10, together=True)
    assert np.isfinite(ll) and len(dists) == 5
    # a copy made before the hook was set trains on
    cp.train(5)
    assert cp.global_step == 10
    mdl._ctx.fail_at_step(-1)
    mdl.train(2)
    assert np.isfinite(mdl.ll_per_word)
