"""The LDAModel surface on the GPU, mirroring tomotopy's test/unit_test.py
LDA cases (null_doc, train1/train4/train0, train0_without_optim,
reproducibility, infer, uninit_doc, asymmetric alpha) on the reference's own
sample.txt.  Iteration counts are reduced from the reference's 2000 to keep
the suite short; the reference's cases are smoke tests plus one
reproducibility assertion, which is kept exactly."""
import os
import warnings

import numpy as np
import pytest

import tomotopy_amd as tp

pytestmark = pytest.mark.gpu

SAMPLE = os.path.join(os.path.dirname(__file__), "golden", "sample.txt")
PARALLEL = [tp.ParallelScheme.COPY_MERGE, tp.ParallelScheme.PARTITION]


def _lines():
    with open(SAMPLE, encoding="utf-8") as f:
        return [l.strip().split() for l in f]


def _model(**kw):
    mdl = tp.LDAModel(**kw)
    for ch in _lines():
        if len(ch) < 1:
            continue
        mdl.add_doc(ch)
    return mdl


@pytest.mark.parametrize("ps", PARALLEL)
def test_null_doc(ps):
    mdl = _model(tw=0, min_df=200, rm_top=200, k=40)
    mdl.train(100, workers=1, parallel=ps)
    assert mdl.num_vocabs == 0
    assert len(mdl.docs[0].words) > 0
    assert all(t == -1 for t in mdl.docs[0].topics)


@pytest.mark.parametrize("ps", PARALLEL)
@pytest.mark.parametrize("workers", [1, 4, 0])
def test_train(ps, workers):
    mdl = _model(tw=0, min_df=2, rm_top=2, k=40)
    mdl.train(200, workers=workers, parallel=ps)
    assert mdl.global_step == 200
    assert -12 < mdl.ll_per_word < -5
    assert sum(mdl.get_count_by_topics()) == mdl.num_words
    mdl.summary(file=open(os.devnull, "w"))


def test_train_without_optim_and_asymmetric_alpha():
    mdl = _model(tw=0, min_df=2, rm_top=2, k=40, alpha=[0.1 * (i + 1) for i in range(40)])
    mdl.optim_interval = 0
    mdl.train(100, workers=0)
    assert np.allclose(mdl.alpha, [0.1 * (i + 1) for i in range(40)], rtol=1e-6)


def test_reproducibility():
    results = []
    for _ in range(3):
        mdl = _model(tw=0, min_df=2, rm_top=2, seed=42, k=40)
        mdl.train(300, workers=1)
        results.append([mdl.get_topic_words(k) for k in range(mdl.k)])
    assert results[0] == results[1]
    assert results[0] == results[2]


def test_infer_and_uninit_doc():
    lines = _lines()
    mdl = tp.LDAModel(tw=0, min_df=2, rm_top=2, k=40)
    for ch in lines[5:]:
        mdl.add_doc(ch)
    mdl.train(50)
    unseen = [mdl.make_doc(ch) for ch in lines[:5]]
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        with pytest.raises(Warning):
            unseen[0].get_topics()
    dists, ll = mdl.infer(unseen)
    assert len(dists) == 5 and len(ll) == 5
    for d in dists:
        assert abs(sum(d) - 1.0) < 1e-4
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        unseen[0].get_topics()
        unseen[0].get_topic_dist()
    single, ll1 = mdl.infer(unseen[1])
    assert np.allclose(single, dists[1]) and ll1 == ll[1]
    with pytest.raises(RuntimeError):
        mdl.add_doc(["new", "words"])


def test_copy_train():
    mdl = _model(tw=0, min_df=2, rm_top=2, k=10)
    mdl.train(50)
    new = mdl.copy()
    del mdl
    new.train(50)
    assert new.global_step == 100


@pytest.mark.parametrize("ps", PARALLEL)
def test_word_prior(ps):
    """unit_test.py word_prior: each document's first word gets the prior
    {n % k: 1} (the first prior set for a word stays), then training and
    summary.  Iterations reduced from 2000."""
    k = 10
    mdl = tp.LDAModel(tw=0, min_df=2, rm_top=2, k=k)
    n_prior = 0
    for n, ch in enumerate(_lines()):
        if len(ch) < 1:
            continue
        mdl.add_doc(ch)
        mdl.set_word_prior(ch[0], {n % k: 1})
        n_prior += 1
    mdl.train(100, workers=1, parallel=ps)
    assert np.isfinite(mdl.ll_per_word)
    mdl.summary(file=open(os.devnull, "w"))


@pytest.mark.parametrize("ps", PARALLEL)
def test_train_multi_corpus(ps):
    """unit_test.py train_multi_corpus with pre-tokenised words: two corpora
    added with add_corpus, trained, and each returned document's
    log-likelihood read after training."""
    lines = [ch for ch in _lines() if ch]
    corpus1, corpus2 = tp.utils.Corpus(), tp.utils.Corpus()
    for i, ch in enumerate(lines):
        (corpus1 if i < 10 else corpus2).add_doc(ch)
    mdl = tp.LDAModel(min_cf=2, rm_top=2, k=10)
    t1 = mdl.add_corpus(corpus1)
    t2 = mdl.add_corpus(corpus2)
    mdl.train(50, parallel=ps)
    assert len(t1) == 10 and len(t2) == len(lines) - 10
    for d in list(t1[:10]) + list(t2[:10]):
        assert np.isfinite(d.get_ll())
        assert abs(sum(d.get_topic_dist()) - 1.0) < 1e-4


def test_train_infer_corpus_split():
    """unit_test.py train_infer_raw_corpus with pre-tokenised words: train on
    four fifths of a Corpus, infer the other fifth (a Corpus slice), read
    each result's log-likelihood and words."""
    corpus = tp.utils.Corpus()
    for ch in _lines():
        if ch:
            corpus.add_doc(ch)
    test_size = len(corpus) // 5
    test_set, train_set = corpus[:test_size], corpus[test_size:]
    mdl = tp.LDAModel(min_cf=2, rm_top=2, k=10, corpus=train_set)
    mdl.train(50)
    mdl.summary(file=open(os.devnull, "w"))
    result, ll = mdl.infer(test_set)
    assert len(result) == test_size and len(ll) == test_size and np.all(np.isfinite(ll))
    for r in result:
        assert np.isfinite(r.get_ll())
        assert r[0] in mdl.vocabs
