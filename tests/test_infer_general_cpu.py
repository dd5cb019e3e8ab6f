"""Fold-in of term-weighted (IDF / PMI) models and of models with word priors
in the oracle (LDAModel.hpp:920-953): the per-document local state against
tomoto's full K x V private copy, the reductions to the integer fold-in, the
initial draw, and theta / log-likelihood against scipy closed forms."""
import numpy as np
import pytest
from scipy.special import gammaln

from oracle import oracle as O
from tomotopy_amd import corpus, synth

ALPHA, ETA = 0.1, 0.01


def _prior_table(V, K, rng, n_prior=12):
    eta_kv = np.full((V, K), np.float32(ETA), np.float32)
    for v in rng.choice(V, size=n_prior, replace=False):
        eta_kv[v] = (rng.random(K) * rng.choice([0.01, 1.0, 5.0])).astype(np.float32)
    return eta_kv


def _setup(kind, K=8, seed=3, n_docs=60):
    """A model trained for 3 sweeps and 6 of its documents (with removed words,
    ids >= V) as the unseen batch.  kind: one, prior, idf, pmi, pmi_prior."""
    words, off, nw, _ = synth.workload("plain_k20", n_docs=n_docs)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw, min_cf=2, rm_top=1)
    rng = np.random.default_rng(seed)
    weighted = kind in ("idf", "pmi", "pmi_prior")
    wts = corpus.term_weights(ws, offs, V, "idf" if kind == "idf" else "pmi") if weighted else None
    m = (O.WeightedModel if weighted else O.Model)(K, V, ALPHA, ETA, seed)
    eta_kv = None
    if kind in ("prior", "pmi_prior"):
        eta_kv = _prior_table(V, K, rng)
        m.set_priors(eta_kv)
    m.set_optim(0, 0)
    if weighted:
        m.set_corpus(ws, offs, wts)
    else:
        m.set_corpus(ws, offs)
    m.train(3, workers=1, scheme=1)
    e = int(offs[6])
    uw, uoff = ws[:e].copy(), offs[:7].copy()
    uwt = None if wts is None else wts[:e].copy()
    return m, V, uw, uoff, uwt, eta_kv


@pytest.mark.parametrize("kind", ["prior", "idf", "pmi", "pmi_prior"])
def test_local_state_equals_full_private_copy(kind):
    """The document-local columns and the changed-terms LL give tomoto's
    full-copy fold-in: same topics, same theta, LL to FP64 rounding."""
    m, V, uw, uoff, uwt, _ = _setup(kind)
    th0, ll0, z0 = m.infer_gen(uw, uoff, 6, infer_seed=5, weights=uwt, faithful=False)
    th1, ll1, z1 = m.infer_gen(uw, uoff, 6, infer_seed=5, weights=uwt, faithful=True)
    assert np.array_equal(z0, z1)
    assert np.array_equal(th0, th1)
    np.testing.assert_allclose(ll0, ll1, rtol=1e-9)


@pytest.mark.parametrize("kind", ["one", "prior", "pmi"])
def test_tomoto_form_with_lgamma_table(kind):
    """faithful=2 (tomoto's full-copy LL with its lgamma table, the form the
    CPU baseline times): the same topics and theta; the LL within the table's
    interpolation error of FP64; for ONE without priors, the integer fold-in's
    table mode (changed cells only, the same table) to summation order."""
    m, V, uw, uoff, uwt, _ = _setup(kind)
    th0, ll0, z0 = m.infer_gen(uw, uoff, 6, infer_seed=5, weights=uwt, faithful=False)
    th2, ll2, z2 = m.infer_gen(uw, uoff, 6, infer_seed=5, weights=uwt, faithful=2)
    assert np.array_equal(z0, z2) and np.array_equal(th0, th2)
    assert not np.array_equal(ll0, ll2)  # the table is in use
    np.testing.assert_allclose(ll2, ll0, rtol=1e-4)
    if kind == "one":
        _, ll3, z3 = m.infer(uw, uoff, 6, infer_seed=5, mode=1)
        assert np.array_equal(z3, z0)
        np.testing.assert_allclose(ll2, ll3, rtol=1e-6)


def test_one_without_priors_is_the_integer_fold_in():
    m, V, uw, uoff, _, _ = _setup("one")
    th0, ll0, z0 = m.infer_gen(uw, uoff, 7, infer_seed=9)
    th1, ll1, z1 = m.infer(uw, uoff, 7, infer_seed=9, mode=0)
    assert np.array_equal(z0, z1)
    assert np.array_equal(th0, th1)
    np.testing.assert_allclose(ll0, ll1, rtol=1e-10)


def test_unit_weights_are_the_integer_fold_in():
    """A weighted model trained with weights 1.0 holds integer counts; its
    fold-in with unit weights takes the integer sampler's path."""
    words, off, nw, _ = synth.workload("plain_k20", n_docs=50)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw, min_cf=2)
    K = 10
    mi = O.Model(K, V, ALPHA, ETA, 4)
    mw = O.WeightedModel(K, V, ALPHA, ETA, 4)
    for mm in (mi, mw):
        mm.set_optim(0, 0)
    mi.set_corpus(ws, offs)
    mw.set_corpus(ws, offs, np.ones(len(ws), np.float32))
    mi.train(3, workers=1, scheme=1)
    mw.train(3, workers=1, scheme=1)
    e = int(offs[5])
    th0, ll0, z0 = mw.infer_gen(ws[:e], offs[:6], 5, infer_seed=2, weights=np.ones(e, np.float32))
    th1, ll1, z1 = mi.infer(ws[:e], offs[:6], 5, infer_seed=2, mode=0)
    assert np.array_equal(z0, z1)
    assert np.array_equal(th0, th1)
    np.testing.assert_allclose(ll0, ll1, rtol=1e-10)


def test_prior_initial_draw_uses_the_eta_column():
    """iterations = 0: every token's topic is sampleFromDiscrete of its word's
    eta column with the uniform keyed by (infer_seed, position, 0, INFINIT)."""
    m, V, uw, uoff, _, eta_kv = _setup("prior", K=11)
    _, _, z = m.infer_gen(uw, uoff, 0, infer_seed=13)
    for d in range(len(uoff) - 1):
        b, e = int(uoff[d]), int(uoff[d + 1])
        for i in range(e - b):
            v = int(uw[b + i])
            if v >= V:
                assert z[b + i] == 0xFFFF
                continue
            u = O.token_uniform(13, i, 0, "infinit")
            assert z[b + i] == O.sample_from_discrete(eta_kv[v], u)


def _closed_form_ll(nkw, nk, eta_kv, es, alpha, words, V, z, w):
    """getLLRest(model + doc) - getLLRest(model) + getLLDocs(doc) in FP64 from
    the topics; w: token weights (the counts are then their sums)."""
    K = len(nk)
    valid = words < V
    v, k, ww = words[valid].astype(np.int64), z[valid].astype(np.int64), w[valid].astype(np.float64)
    nd = np.bincount(k, weights=ww, minlength=K)
    c = np.zeros((V, K))
    np.add.at(c, (v, k), ww)
    touched = np.unique(v)
    nkw64, nk64, e64 = nkw.astype(np.float64), nk.astype(np.float64), eta_kv.astype(np.float64)
    ll = np.sum(gammaln(nk64 + es) - gammaln(nk64 + nd + es))
    g0, g1, ee = nkw64[touched], nkw64[touched] + c[touched], e64[touched]
    t1 = np.where(g1 != 0, gammaln(g1 + ee) - gammaln(ee), 0.0)
    t0 = np.where(g0 != 0, gammaln(g0 + ee) - gammaln(ee), 0.0)
    ll += np.sum(t1 - t0)
    a = alpha.astype(np.float64)
    ll -= gammaln(ww.sum() + a.sum()) - gammaln(a.sum())
    ll += np.sum(np.where(nd != 0, gammaln(nd + a) - gammaln(a), 0.0))
    return ll, nd


def test_prior_fold_in_closed_forms():
    """Integer counts with word priors: LL with eta_kv and sum_v eta_kv (the
    FP64 sums), theta = (n_dk + a) / (L + sum a), both from the topics."""
    K = 9
    m, V, uw, uoff, _, eta_kv = _setup("prior", K=K)
    z_m, _, nkw, nk = m.state()
    th, ll, z = m.infer_gen(uw, uoff, 8, infer_seed=21)
    es = eta_kv.astype(np.float64).sum(axis=0)
    alpha = np.full(K, np.float32(ALPHA))
    for d in range(len(uoff) - 1):
        b, e = int(uoff[d]), int(uoff[d + 1])
        ref, nd = _closed_form_ll(nkw, nk, eta_kv, es, alpha, uw[b:e], V, z[b:e], np.ones(e - b))
        assert abs(ll[d] - ref) <= 1e-9 * max(1.0, abs(ref))
        L = int((uw[b:e] < V).sum())
        np.testing.assert_array_equal(th[d], (nd.astype(np.float32) + alpha) / (np.float32(L) + np.float32(ALPHA * K)))


@pytest.mark.parametrize("kind", ["idf", "pmi"])
def test_weighted_fold_in_closed_forms(kind):
    """Float counts: the running sums equal the weight sums of the topics up
    to float rounding, so theta and the LL match the closed forms closely; a
    wrong weight, a dropped clamp or an unweighted length moves them by far
    more."""
    K = 8
    m, V, uw, uoff, uwt, _ = _setup(kind, K=K)
    _, _, nkw, nk, _ = m.state()
    th, ll, z = m.infer_gen(uw, uoff, 8, infer_seed=4, weights=uwt)
    eta_kv = np.full((V, K), np.float32(ETA), np.float32)
    es = np.full(K, V * np.float64(np.float32(ETA)))
    alpha = np.full(K, np.float32(ALPHA))
    for d in range(len(uoff) - 1):
        b, e = int(uoff[d]), int(uoff[d + 1])
        ref, nd = _closed_form_ll(nkw, nk, eta_kv, es, alpha, uw[b:e], V, z[b:e], uwt[b:e])
        # float n_k of a few thousand drifts by ~1e-3 over the sweeps; the LL
        # moves by psi(n_k) times that per topic
        assert abs(ll[d] - ref) <= 1e-4 * max(1.0, abs(ref))
        sw = float(uwt[b:e][uw[b:e] < V].astype(np.float64).sum())
        # an emptied topic keeps a rounding-level residue of the float n_dk
        np.testing.assert_allclose(th[d], (nd + ALPHA) / (sw + ALPHA * K), rtol=1e-5, atol=1e-6)


# ------------------------------------------------------------ together=True --
def _unseen(V, n=10, seed=8):
    uw, uoff = synth.unseen_docs(seed, n, V, 20, np.log(60.0), 1.0, 5, 300, zipf_s=None)
    return corpus.csr_from_arrays(uw, uoff, np.arange(V))


@pytest.mark.parametrize("scheme,workers", [(1, 1), (3, 3), (2, 2)])
def test_together_prior_closed_forms(scheme, workers):
    """together=True with word priors: the initial topics come from the eta
    columns, and from the returned topics theta = (n_dk+a)/(L+sum a) and
    ll = getLLRest(model + docs) - getLLRest(model) + getLLDocs(docs) with
    eta_kv and sum_v eta_kv, by scipy."""
    K = 9
    m, V, _, _, _, eta_kv = _setup("prior", K=K)
    _, _, nkw0, nk0 = m.state()
    uw, uoff = _unseen(V)
    theta, ll, z = m.infer_together(uw, uoff, 6, infer_seed=19, workers=workers, scheme=scheme)
    D = len(uoff) - 1
    ndk = np.zeros((D, K), np.int64)
    nkw = nkw0.astype(np.int64).copy()
    for d in range(D):
        for g in range(int(uoff[d]), int(uoff[d + 1])):
            if uw[g] < V:
                ndk[d, z[g]] += 1
                nkw[uw[g], z[g]] += 1
    nk = nkw.sum(axis=0)
    e = eta_kv.astype(np.float64)
    es = e.sum(axis=0)

    def rest(nkw_, nk_):
        return np.sum(gammaln(es) - gammaln(nk_ + es)) + np.sum(np.where(nkw_ != 0, gammaln(nkw_ + e) - gammaln(e), 0.0))
    a = np.full(K, np.float64(np.float32(ALPHA)))
    L = ndk.sum(axis=1)
    docs = np.sum(gammaln(a.sum()) - gammaln(L + a.sum())) + np.sum(gammaln(ndk + a) - gammaln(a))
    ref = rest(nkw, nk) - rest(nkw0.astype(np.int64), nk0.astype(np.int64)) + docs
    assert ll == pytest.approx(ref, rel=1e-9)
    asum = np.float32(0)
    for _ in range(K):
        asum = np.float32(asum + np.float32(ALPHA))
    th = ((ndk.astype(np.float32) + np.float32(ALPHA)) / (L.astype(np.float32)[:, None] + asum)).astype(np.float32)
    assert np.array_equal(theta, th)


def test_together_prior_initial_draw():
    m, V, _, _, _, eta_kv = _setup("prior", K=11)
    uw, uoff = _unseen(V, n=6)
    _, _, z = m.infer_together(uw, uoff, 0, infer_seed=23)
    for g in range(len(uw)):
        if uw[g] < V:
            assert z[g] == O.sample_from_discrete(eta_kv[uw[g]], O.token_uniform(23, g, 0, "init"))


@pytest.mark.parametrize("scheme,workers", [(1, 1), (3, 3), (2, 2)])
def test_together_unit_weights_are_the_integer_together(scheme, workers):
    words, off, nw, _ = synth.workload("plain_k20", n_docs=50)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw, min_cf=2)
    K = 10
    mi = O.Model(K, V, ALPHA, ETA, 4)
    mw = O.WeightedModel(K, V, ALPHA, ETA, 4)
    for mm in (mi, mw):
        mm.set_optim(0, 0)
    mi.set_corpus(ws, offs)
    mw.set_corpus(ws, offs, np.ones(len(ws), np.float32))
    mi.train(3, workers=1, scheme=1)
    mw.train(3, workers=1, scheme=1)
    uw, uoff = _unseen(V)
    th0, ll0, z0 = mw.infer_together(uw, uoff, 5, infer_seed=2, workers=workers, scheme=scheme,
                                     weights=np.ones(len(uw), np.float32))
    th1, ll1, z1 = mi.infer_together(uw, uoff, 5, infer_seed=2, workers=workers, scheme=scheme)
    assert np.array_equal(z0, z1)
    assert np.array_equal(th0, th1)
    assert ll0 == pytest.approx(ll1, rel=1e-12)


@pytest.mark.parametrize("kind", ["idf", "pmi", "pmi_prior"])
def test_together_weighted_closed_forms(kind):
    """Float counts: theta and the LL match the closed forms over the weight
    sums of the returned topics to float rounding."""
    K = 8
    m, V, _, _, _, eta_kv = _setup(kind, K=K)
    _, _, nkw0, nk0, _ = m.state()
    uw, uoff = _unseen(V)
    uwt = corpus.term_weights(uw, uoff, V, "idf" if kind == "idf" else "pmi",
                              vocab_w=np.linspace(0.5, 2.0, V).astype(np.float32))
    theta, ll, z = m.infer_together(uw, uoff, 6, infer_seed=3, workers=2, scheme=3, weights=uwt)
    D = len(uoff) - 1
    valid = uw < V
    doc = np.repeat(np.arange(D), np.diff(uoff.astype(np.int64)))
    ndk = np.zeros((D, K))
    np.add.at(ndk, (doc[valid], z[valid].astype(np.int64)), uwt[valid].astype(np.float64))
    nkw = nkw0.astype(np.float64).copy()
    np.add.at(nkw, (uw[valid].astype(np.int64), z[valid].astype(np.int64)), uwt[valid].astype(np.float64))
    e = np.full((V, K), np.float64(np.float32(ETA))) if eta_kv is None else eta_kv.astype(np.float64)
    es = np.full(K, V * np.float64(np.float32(ETA))) if eta_kv is None else e.sum(axis=0)

    def rest(nkw_):
        nk_ = nkw_.sum(axis=0)
        return np.sum(gammaln(es) - gammaln(nk_ + es)) + np.sum(np.where(nkw_ != 0, gammaln(nkw_ + e) - gammaln(e), 0.0))
    a = np.full(K, np.float64(np.float32(ALPHA)))
    sw = np.zeros(D)
    np.add.at(sw, doc[valid], uwt[valid].astype(np.float64))
    docs = np.sum(gammaln(a.sum()) - gammaln(sw + a.sum())) + np.sum(np.where(ndk != 0, gammaln(ndk + a) - gammaln(a), 0.0))
    ref = rest(nkw) - rest(nkw0.astype(np.float64)) + docs
    assert abs(ll - ref) <= 1e-4 * abs(ref)
    np.testing.assert_allclose(theta, (ndk + ALPHA) / (sw[:, None] + ALPHA * K), rtol=1e-5, atol=1e-6)


# ----------------------------------------------------------- segmented ------
@pytest.mark.parametrize("T,M", [(40, 8), (100, 2), (7, 4)])
def test_segmented_one_is_the_integer_segmented_fold_in(T, M):
    """Integer counts without priors: the segmented rule on running counts
    (merged once per iteration) is the integer fold-in's segmented rule with
    per-iteration merges (pinned by its own tests)."""
    m, V, uw, uoff, _, _ = _setup("one")
    th0, ll0, z0 = m.infer_gen(uw, uoff, 6, infer_seed=3, seg_tokens=T, seg_max=M)
    th1, ll1, z1 = m.infer(uw, uoff, 6, infer_seed=3, mode=0, seg_tokens=T, seg_max=M, seg_sync=0)
    assert np.array_equal(z0, z1)
    assert np.array_equal(th0, th1)
    np.testing.assert_allclose(ll0, ll1, rtol=1e-10)
    # and it does split: the sequential sampler gives other topics
    _, _, zs = m.infer_gen(uw, uoff, 6, infer_seed=3)
    assert not np.array_equal(z0, zs)


def test_segmented_unit_weights_are_the_integer_segmented_fold_in():
    words, off, nw, _ = synth.workload("plain_k20", n_docs=50)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw, min_cf=2)
    K = 10
    mi = O.Model(K, V, ALPHA, ETA, 4)
    mw = O.WeightedModel(K, V, ALPHA, ETA, 4)
    for mm in (mi, mw):
        mm.set_optim(0, 0)
    mi.set_corpus(ws, offs)
    mw.set_corpus(ws, offs, np.ones(len(ws), np.float32))
    mi.train(3, workers=1, scheme=1)
    mw.train(3, workers=1, scheme=1)
    e = int(offs[6])
    th0, ll0, z0 = mw.infer_gen(ws[:e], offs[:7], 5, infer_seed=2, weights=np.ones(e, np.float32), seg_tokens=30,
                                seg_max=8)
    th1, ll1, z1 = mi.infer_gen(ws[:e], offs[:7], 5, infer_seed=2, seg_tokens=30, seg_max=8)
    assert np.array_equal(z0, z1)
    assert np.array_equal(th0, th1)
    np.testing.assert_allclose(ll0, ll1, rtol=1e-10)


@pytest.mark.parametrize("kind", ["prior", "pmi", "pmi_prior"])
def test_segmented_general_closed_forms(kind):
    """Segmented fold-in of prior / PMI models: theta and the LL against the
    closed forms from the returned topics (integer counts exactly; float
    counts to their rounding)."""
    K = 8
    m, V, uw, uoff, uwt, eta_kv = _setup(kind, K=K)
    st = m.state()
    nkw, nk = st[2], st[3]
    th, ll, z = m.infer_gen(uw, uoff, 8, infer_seed=4, weights=uwt, seg_tokens=25, seg_max=4)
    e_kv = np.full((V, K), np.float32(ETA), np.float32) if eta_kv is None else eta_kv
    es = (np.full(K, V * np.float64(np.float32(ETA))) if eta_kv is None else eta_kv.astype(np.float64).sum(axis=0))
    alpha = np.full(K, np.float32(ALPHA))
    for d in range(len(uoff) - 1):
        b, e = int(uoff[d]), int(uoff[d + 1])
        w = np.ones(e - b) if uwt is None else uwt[b:e]
        ref, nd = _closed_form_ll(nkw, nk, e_kv, es, alpha, uw[b:e], V, z[b:e], w)
        tol = 1e-9 if uwt is None else 1e-4
        assert abs(ll[d] - ref) <= tol * max(1.0, abs(ref))
        sw = float(np.asarray(w, np.float64)[uw[b:e] < V].sum())
        np.testing.assert_allclose(th[d], (nd + ALPHA) / (sw + ALPHA * K), rtol=1e-5, atol=1e-6)
