"""Fold-in of term-weighted (IDF / PMI) models and of models with word priors
on the device against the CPU oracle: topics and theta bit-exact, FP64
log-likelihood to 1e-9, batched = plain, with empty documents and documents
of removed words only; and the LDAModel.infer surface of those models."""
import numpy as np
import pytest

from oracle import oracle as O
import tomotopy_amd as tp
from tomotopy_amd import _native, corpus, synth

pytestmark = pytest.mark.gpu

TW = {"one": 0, "prior": 0, "idf": 1, "pmi": 2, "pmi_prior": 2}


def _setup(kind, K, seed, n_docs=80):
    """Device context and oracle model trained identically (none scheme, which
    the schedule tests hold bit-exact), and an unseen batch: 7 training
    documents, an empty one and one of removed words only."""
    words, off, nw, _ = synth.workload("plain_k20", n_docs=n_docs)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw, min_cf=2, rm_top=1)
    rng = np.random.default_rng(seed)
    tw = TW[kind]
    wts = corpus.term_weights(ws, offs, V, "pmi" if tw == 2 else "idf") if tw else None
    ctx = _native.Context(K, V, 0.1, 0.01, seed=seed, tw=tw)
    m = (O.WeightedModel if tw else O.Model)(K, V, 0.1, 0.01, seed, scan_order=1)
    if kind.endswith("prior"):
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
    e = int(offs[7])
    n_rm = int(np.count_nonzero(ws >= V))
    assert n_rm > 0
    uw = np.concatenate([ws[:e], np.array([V, V], np.uint32)])
    uoff = np.concatenate([offs[:8], [e, e + 2]]).astype(np.uint64)
    uwt = None
    if wts is not None:
        uwt = np.concatenate([wts[:e], np.zeros(2, np.float32)])
    return ctx, m, uw, uoff, uwt, V


@pytest.mark.parametrize("kind,K", [("prior", 20), ("prior", 200), ("idf", 20), ("idf", 1000), ("pmi", 200),
                                    ("pmi_prior", 20), ("pmi_prior", 500)])
def test_fold_in_bit_exact(kind, K):
    ctx, m, uw, uoff, uwt, V = _setup(kind, K, seed=K + len(kind))
    print(kind, "K", K, flush=True)
    tho, llo, zo = m.infer_gen(uw, uoff, 12, infer_seed=77, weights=uwt)
    for mode in (0, 1):
        thg, llg, zg = ctx.infer(uw, uoff, 12, infer_seed=77, mode=mode, want_z=True, weights=uwt)
        assert np.array_equal(zg, zo), "mode %d: z differs at %d tokens" % (mode, int((zg != zo).sum()))
        assert np.array_equal(thg, tho), "mode %d theta" % mode
        np.testing.assert_allclose(llg, llo, rtol=1e-9, atol=1e-9)
    # the empty document: theta = alpha / sum(alpha) (the sum in float), ll = 0
    np.testing.assert_allclose(thg[7], np.full(K, 1.0 / K, np.float32), rtol=1e-5)
    assert llg[7] == 0.0 and llg[8] == 0.0


def test_term_weighted_fold_in_needs_weights():
    ctx, m, uw, uoff, uwt, V = _setup("idf", 20, seed=1)
    with pytest.raises(ValueError):
        ctx.infer(uw, uoff, 5)
    bad = uwt.copy()
    bad[3] = np.nan
    with pytest.raises(ValueError):
        ctx.infer(uw, uoff, 5, weights=bad)
    bad[3] = -1.0
    with pytest.raises(ValueError):
        ctx.infer(uw, uoff, 5, weights=bad)
    with pytest.raises(ValueError):
        ctx.infer_together(uw, uoff, 5)


@pytest.mark.parametrize("tw", [tp.TermWeight.IDF, tp.TermWeight.PMI])
def test_model_surface_infer_term_weighted(tw):
    """LDAModel.infer of an IDF / PMI model: the token weights come from the
    model's vocabulary weights (LDAModel.hpp:794-809), theta sums to one,
    a document inferred alone equals the same document in a batch."""
    words, off, nw, _ = synth.workload("plain_k20", n_docs=120)
    vocab = ["w%d" % i for i in range(nw)]
    mdl = tp.LDAModel(tw=tw, k=10, min_cf=2, seed=4)
    for d in range(len(off) - 1):
        mdl.add_doc([vocab[int(w)] for w in words[int(off[d]):int(off[d + 1])]])
    mdl.train(10, workers=1)
    docs = [mdl.make_doc([vocab[int(w)] for w in words[int(off[d]):int(off[d + 1])]]) for d in range(5)]
    dists, ll = mdl.infer(docs, iterations=20)
    assert len(dists) == 5 and np.all(np.isfinite(ll))
    for t in dists:
        assert abs(sum(t) - 1.0) < 1e-5
    one, ll1 = mdl.infer(mdl.make_doc([vocab[int(w)] for w in words[int(off[2]):int(off[3])]]), iterations=20)
    assert np.array_equal(np.asarray(one, np.float32), np.asarray(dists[2], np.float32)) and ll1 == ll[2]


def test_model_surface_infer_with_priors():
    words, off, nw, _ = synth.workload("plain_k20", n_docs=100)
    vocab = ["w%d" % i for i in range(nw)]
    mdl = tp.LDAModel(k=6, seed=2)
    for d in range(len(off) - 1):
        mdl.add_doc([vocab[int(w)] for w in words[int(off[d]):int(off[d + 1])]])
    mdl.set_word_prior(vocab[int(words[0])], [5.0, 0.01, 0.01, 0.01, 0.01, 0.01])
    mdl.train(10, workers=1)
    doc = mdl.make_doc([vocab[int(w)] for w in words[int(off[0]):int(off[1])]])
    dist, ll = mdl.infer(doc, iterations=20)
    assert abs(sum(dist) - 1.0) < 1e-5 and np.isfinite(ll)


@pytest.mark.parametrize("kind,K,T,M", [("prior", 20, 30, 8), ("idf", 200, 20, 8), ("pmi", 20, 12, 2),
                                        ("pmi_prior", 500, 25, 8), ("idf", 1000, 40, 4)])
def test_segmented_fold_in_bit_exact(kind, K, T, M):
    """Documents longer than T tokens split into up to M wavefronts: topics,
    theta bit-exact and LL to 1e-9 against the oracle's segmented rule,
    batched (segment-count buckets on their own streams) and plain."""
    ctx, m, uw, uoff, uwt, V = _setup(kind, K, seed=K + 3)
    ctx.set_infer_options(T, M)
    print(kind, "K", K, "seg", T, M, flush=True)
    tho, llo, zo = m.infer_gen(uw, uoff, 9, infer_seed=31, weights=uwt, seg_tokens=T, seg_max=M)
    _, _, zs = m.infer_gen(uw, uoff, 9, infer_seed=31, weights=uwt)
    assert not np.array_equal(zo, zs)  # the rule does split these documents
    for mode in (0, 1):
        thg, llg, zg = ctx.infer(uw, uoff, 9, infer_seed=31, mode=mode, want_z=True, weights=uwt)
        assert np.array_equal(zg, zo), "mode %d: z differs at %d tokens" % (mode, int((zg != zo).sum()))
        assert np.array_equal(thg, tho), "mode %d theta" % mode
        np.testing.assert_allclose(llg, llo, rtol=1e-9, atol=1e-9)
