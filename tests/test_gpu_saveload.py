"""LDAModel.save / saves / load / loads on the GPU, after the reference's
save_and_load and test_issue197 cases (test/unit_test.py:208-226, 298-304):
the loaded model carries the same counts, topics, hyperparameters and
vocabulary, trains on, and re-saves to identical bytes."""
import os

import numpy as np
import pytest

import tomotopy_amd as tp

pytestmark = pytest.mark.gpu

SAMPLE = os.path.join(os.path.dirname(__file__), "golden", "sample.txt")


def _model(n_lines=None, **kw):
    mdl = tp.LDAModel(**kw)
    with open(SAMPLE, encoding="utf-8") as f:
        for i, line in enumerate(f):
            if n_lines is not None and i >= n_lines:
                break
            ch = line.strip().split()
            if ch:
                mdl.add_doc(ch)
    return mdl


def _snapshot(mdl):
    return {"z": mdl._state("z").copy(), "nkw": mdl._state("nkw").copy(), "nk": mdl._state("nk").copy(),
            "ndk": mdl._state("ndk").copy(), "alpha": list(mdl.alpha), "step": mdl.global_step,
            "vocab": list(mdl.used_vocabs), "llpw": mdl.ll_per_word}


def _assert_same(a, b, tw):
    for k in ("nkw", "nk"):
        assert np.array_equal(a[k], b[k]), k
    if tw == 0:
        assert np.array_equal(a["ndk"], b["ndk"])
        assert a["llpw"] == pytest.approx(b["llpw"], rel=1e-12)
    else:
        # n_dk is rebuilt from the topics on load (DocumentLDA::update,
        # LDAModel.hpp:1163-1173): float sums that differ in the last bits
        # from the ones training accumulated
        np.testing.assert_allclose(a["ndk"], b["ndk"], rtol=1e-4, atol=1e-3)
        assert a["llpw"] == pytest.approx(b["llpw"], rel=1e-6)
    assert a["alpha"] == b["alpha"] and a["step"] == b["step"] and a["vocab"] == b["vocab"]


@pytest.mark.parametrize("tw", [0, 1, 2])
def test_save_and_load(tw, tmp_path):
    mdl = _model(n_lines=300, tw=tw, min_df=2, rm_top=2, k=10, seed=3)
    mdl.train(20, parallel=tp.ParallelScheme.PARTITION)
    before = _snapshot(mdl)
    path = str(tmp_path / "m.bin")
    mdl.save(path)
    print("tw", tw, "file bytes", os.path.getsize(path), flush=True)
    m2 = tp.LDAModel.load(path)
    assert m2.tw == tw and m2.k == 10 and m2.num_words == mdl.num_words
    assert m2.removed_top_words == mdl.removed_top_words
    _assert_same(before, _snapshot(m2), tw)
    valid = mdl._words < mdl._real_v
    assert np.array_equal(_snapshot(m2)["z"][valid], before["z"][valid])
    assert [d.words for d in m2.docs[:5]] == [d.words for d in mdl.docs[:5]]
    m2.train(20, parallel=tp.ParallelScheme.PARTITION)
    assert m2.global_step == 40
    b = m2.saves()
    m3 = tp.LDAModel.loads(b)
    assert m3.saves() == b  # re-saving a loaded model reproduces the bytes
    m3.train(20, parallel=tp.ParallelScheme.PARTITION)
    assert np.isfinite(m3.ll_per_word)


def test_resume_is_deterministic():
    """The none scheme with a fixed seed: 10 iterations, save, load, 10 more
    equals 20 uninterrupted iterations (the sampler's uniforms are keyed by
    seed, token and iteration; alpha and the counts travel in the file)."""
    a = _model(n_lines=200, k=8, seed=11)
    a.optim_interval = 5
    a.train(20, workers=1)
    b = _model(n_lines=200, k=8, seed=11)
    b.optim_interval = 5
    b.train(10, workers=1)
    c = tp.LDAModel.loads(b.saves())
    c.train(10, workers=1)
    assert np.array_equal(a._state("z"), c._state("z"))
    assert np.array_equal(a._state("nkw"), c._state("nkw"))
    assert a.alpha == c.alpha


def test_topic_only_model_infers():
    mdl = _model(n_lines=200, k=6, seed=5)
    mdl.train(10)
    m2 = tp.LDAModel.loads(mdl.saves(full=False))
    assert len(m2.docs) == 0
    for t in range(6):
        assert m2.get_topic_word_dist(t) == mdl.get_topic_word_dist(t)
    doc = m2.make_doc(mdl.docs[0].words)
    dist, ll = m2.infer(doc, iterations=20)
    assert len(dist) == 6 and abs(sum(dist) - 1.0) < 1e-4 and np.isfinite(ll)


def test_issue197_empty_model(tmp_path):
    model = tp.LDAModel()
    path = str(tmp_path / "model.bin")
    model.save(path)
    model = tp.LDAModel.load(path)
    doc = model.make_doc(["unseen", "document"])
    model.infer(doc)


def test_load_rejects_other_files(tmp_path):
    p = tmp_path / "x.bin"
    p.write_bytes(b"not a model")
    with pytest.raises(RuntimeError):
        tp.LDAModel.load(str(p))
