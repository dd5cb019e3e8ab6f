"""Read-outs on the device against closed forms from the downloaded counts:
get_topic_word_dist ((n_kw + eta) / (n_k + V*eta) in float with V*eta formed
as float(V) * eta, LDAModel.hpp:828-840), get_count_by_topics (valid tokens
per topic, 815-826) and Document.get_topic_dist of training documents
((n_dk + alpha) / (weight sum + sum alpha), 1145-1159: the valid token count
for TermWeight ONE, the float sum of the token weights in token order for
IDF / PMI), for integer and float counts, and on a 100,000-word K=1000 model
where only one topic's row leaves the device."""
import os

import numpy as np
import pytest

import tomotopy_amd as tp
from tomotopy_amd import _native, corpus, synth

pytestmark = pytest.mark.gpu

SAMPLE = os.path.join(os.path.dirname(__file__), "golden", "sample.txt")


def _model(tw):
    mdl = tp.LDAModel(tw=tw, min_df=2, rm_top=2, k=15, seed=4)
    with open(SAMPLE, encoding="utf-8") as f:
        for line in f:
            ch = line.strip().split()
            if ch:
                mdl.add_doc(ch)
    mdl.train(20, workers=4, parallel=tp.ParallelScheme.PARTITION)
    return mdl


def _seq_sum(xs):
    acc = np.float32(0.0)
    for x in xs:
        acc = np.float32(acc + np.float32(x))
    return acc


@pytest.mark.parametrize("tw", [0, 2])
def test_readouts_closed_forms(tw):
    mdl = _model(tw)
    K, V = mdl.k, mdl._real_v
    if tw:
        z, ndk, nkw, nk = mdl._ctx.download_state_f()
    else:
        z, ndk, nkw, nk = mdl._ctx.download_state()
    eta = np.float32(mdl.eta)
    veta = np.float32(np.float32(V) * eta)
    for k in range(K):
        col = nkw[:, k].astype(np.float32) + eta
        want = col / np.float32(np.float32(nk[k]) + veta)
        got = np.asarray(mdl.get_topic_word_dist(k), np.float32)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), k
        assert np.array_equal(np.asarray(mdl.get_topic_word_dist(k, normalize=False), np.float32), col)
        top = mdl.get_topic_words(k, top_n=7, return_id=True)
        assert [i for _, i, _ in top] == list(np.argsort(-want, kind="stable")[:7])
    valid = mdl._words < V
    counts = np.bincount(z[valid].astype(np.int64), minlength=K)
    assert mdl.get_count_by_topics() == counts.tolist()
    if tw == 0:
        assert counts.tolist() == nk.astype(np.int64).tolist()
    alpha = np.asarray(mdl.alpha, np.float32)
    asum = _seq_sum(alpha)
    for d in range(0, len(mdl.docs), 5):
        b, e = int(mdl._off[d]), int(mdl._off[d + 1])
        if tw:
            wsum = _seq_sum([w for v, w in zip(mdl._words[b:e], mdl._weights[b:e]) if v < V])
        else:
            wsum = np.float32(int(np.count_nonzero(mdl._words[b:e] < V)))
        num = ndk[d].astype(np.float32) + alpha
        want = num / np.float32(wsum + asum)
        got = np.asarray(mdl.docs[d].get_topic_dist(), np.float32)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), d
        assert np.array_equal(np.asarray(mdl.docs[d].get_topic_dist(normalize=False), np.float32), num)


def test_topic_row_readout_at_k1000():
    words, off, nw, _ = synth.workload("large_k1000", n_docs=3000)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw)
    K = 1000
    ctx = _native.Context(K, V, 0.05, 0.01, seed=9)
    ctx.upload_corpus(ws, offs)
    ctx.train(3, workers=0)
    z, _, nkw, nk = ctx.download_state(ndk=False)
    veta = np.float32(np.float32(V) * np.float32(0.01))
    for k in (0, 1, 499, 998, 999):
        want = (nkw[:, k].astype(np.float32) + np.float32(0.01)) / np.float32(np.float32(nk[k]) + veta)
        got = ctx.topic_word_dist(k)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), k
    cnt = ctx.count_by_topic()
    assert np.array_equal(cnt.astype(np.int64), np.bincount(z[ws < V].astype(np.int64), minlength=K))
    with pytest.raises(ValueError):
        ctx.topic_word_dist(K)
