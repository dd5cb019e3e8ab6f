"""Seeded synthetic corpora from the LDA generative process.

Word ids are dictionary ids in order of first appearance, as tomotopy's
add_doc assigns them (Dictionary::add), so `corpus.prepare_arrays` applies the
same vocabulary ordering a tomotopy user would get.

Workloads (the sizes the engine is measured and checked on):
  plain_k20      1,000 docs, V=5,000, 20 true topics, alpha=0.1, beta=0.01,
                 lengths LogNormal(ln 200, 0.5) clipped to [20, 2000]
  readme_k100    1,000 docs, mean length 1,500, V=50,000, Zipf(1.07)
                 background mixed 50/50 into 100 Dir(0.01) topics
  large_k1000    100,000 docs, mean length 300, V=100,000, Zipf(1.07), 1,000
                 topics
  foldin_k200    20,000 unseen docs, lengths LogNormal(ln 200, 1.2) clipped
                 to [5, 20000], against a K=200, V=50,000 model
  serve_k500     requests of 1..64 docs, lengths LogNormal(ln 150, 1.0)
                 clipped to [5, 10000], against a K=500, V=100,000 model
"""
import numpy as np


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def lognormal_lengths(rng, n, mu, sigma, lo, hi):
    x = np.rint(rng.lognormal(mu, sigma, size=n)).astype(np.int64)
    return np.clip(x, lo, hi)


def _sparse_topic(rng, V, beta, keep_mass=1.0 - 1e-7, max_words=20000):
    g = rng.gamma(beta, size=V)
    s = g.sum()
    if s <= 0:
        g = np.ones(V)
        s = float(V)
    p = g / s
    cand = np.nonzero(p > 1e-12)[0]  # the rest carries < 1e-7 of the mass
    idx = cand[np.argsort(-p[cand], kind="stable")]
    c = np.cumsum(p[idx])
    n = int(min(np.searchsorted(c, keep_mass) + 1, max_words, len(idx)))
    ids = idx[:n]
    pp = p[ids] / p[ids].sum()
    return ids, np.cumsum(pp)


def _topics(rng, V, true_k, beta, zipf_s):
    topics = [_sparse_topic(rng, V, beta) for _ in range(true_k)]
    zcdf = None
    if zipf_s is not None:
        zp = np.arange(1, V + 1, dtype=np.float64) ** (-float(zipf_s))
        zcdf = np.cumsum(zp / zp.sum())
    return topics, zcdf


def _draw(rng, topics, zcdf, V, true_k, alpha, lengths, mix=0.5):
    """Raw word ids (0..V-1) of documents with the given lengths."""
    lengths = np.asarray(lengths, dtype=np.int64)
    n_docs = len(lengths)
    N = int(lengths.sum())
    z = np.empty(N, dtype=np.int32)
    off = np.zeros(n_docs + 1, dtype=np.int64)
    np.cumsum(lengths, out=off[1:])
    chunk = max(1, min(n_docs, 4_000_000 // max(true_k, 1)))
    for d0 in range(0, n_docs, chunk):
        d1 = min(n_docs, d0 + chunk)
        th = rng.gamma(alpha, size=(d1 - d0, true_k))
        s = th.sum(axis=1, keepdims=True)
        th = np.where(s > 0, th / np.where(s > 0, s, 1.0), 1.0 / true_k)
        cdf = np.cumsum(th, axis=1)
        cdf[:, -1] = 1.0
        # row-wise inverse CDF in one searchsorted: row r's cdf lives in [r, r+1]
        rows = np.repeat(np.arange(d1 - d0), lengths[d0:d1])
        u = rng.random(len(rows)) + rows
        flat = (cdf + np.arange(d1 - d0)[:, None]).ravel()
        idx = np.searchsorted(flat, u, side="right")
        z[off[d0]:off[d1]] = np.minimum(idx - rows * true_k, true_k - 1)
    words = np.empty(N, dtype=np.int64)
    if zcdf is not None:
        bg = rng.random(N) < mix
        nb = int(bg.sum())
        words[bg] = np.minimum(np.searchsorted(zcdf, rng.random(nb), side="right"), V - 1)
        tsel = np.nonzero(~bg)[0]
    else:
        tsel = np.arange(N)
    if len(tsel):
        zt = z[tsel]
        order = np.argsort(zt, kind="stable")
        bounds = np.searchsorted(zt[order], np.arange(true_k + 1))
        for k in range(true_k):
            sel = tsel[order[bounds[k]:bounds[k + 1]]]
            if not len(sel):
                continue
            ids, cdf = topics[k]
            words[sel] = ids[np.minimum(np.searchsorted(cdf, rng.random(len(sel)), side="right"), len(ids) - 1)]
    return words, off.astype(np.uint64)


def _first_appearance_ids(words, n_ids):
    """Dictionary ids in order of first appearance; returns (dense words, raw->id map with -1 for unseen)."""
    uniq, first = np.unique(words, return_index=True)
    rank = np.empty(len(uniq), dtype=np.int64)
    rank[np.argsort(first, kind="stable")] = np.arange(len(uniq))
    dense = rank[np.searchsorted(uniq, words)]
    mapping = np.full(n_ids, -1, dtype=np.int64)
    mapping[uniq] = rank
    return dense, mapping


def generate(seed, n_docs, V, true_k, alpha, beta, lengths, zipf_s=None, mix=0.5):
    """Draw a corpus; returns (words, doc_off, n_words) with first-appearance ids."""
    rng = _rng(seed)
    topics, zcdf = _topics(rng, V, true_k, beta, zipf_s)
    words, off = _draw(rng, topics, zcdf, V, true_k, alpha, lengths, mix)
    dense, mapping = _first_appearance_ids(words, V)
    return dense.astype(np.uint32), off, int((mapping >= 0).sum())


def foldin_workload(train_docs=1000, unseen_docs=20000, seed=1004):
    """Fold-in workload: a README-scale training corpus (1,000 docs, mean length
    1,500, V=50,000, Zipf(1.07) background mixed into 100 Dir(0.01) topics) and
    unseen documents with lengths LogNormal(ln 200, 1.2) clipped to [5, 20000]
    drawn from the same topics.  Unseen words are given dictionary ids of the
    training corpus; words the training corpus never saw are dropped, as
    make_doc drops out-of-dictionary words.
    Returns (train_words, train_off, n_words), (unseen_words, unseen_off)."""
    V, true_k = 50000, 100
    rng = _rng(seed)
    topics, zcdf = _topics(rng, V, true_k, 0.01, 1.07)
    tl = lognormal_lengths(_rng(seed + 7919), train_docs, np.log(1500.0) - 0.125, 0.5, 10, 15000)
    tw, toff = _draw(rng, topics, zcdf, V, true_k, 0.1, tl)
    dense, mapping = _first_appearance_ids(tw, V)
    ul = lognormal_lengths(_rng(seed + 104729), unseen_docs, np.log(200.0), 1.2, 5, 20000)
    uw, uoff = _draw(rng, topics, zcdf, V, true_k, 0.1, ul)
    m = mapping[uw]
    keep = m >= 0
    doc_id = np.repeat(np.arange(unseen_docs), np.diff(uoff).astype(np.int64))
    new_len = np.bincount(doc_id[keep], minlength=unseen_docs)
    uoff2 = np.zeros(unseen_docs + 1, dtype=np.uint64)
    np.cumsum(new_len, out=uoff2[1:])
    return (dense.astype(np.uint32), toff, int((mapping >= 0).sum())), (m[keep].astype(np.uint32), uoff2)


def serve_workload(n_requests, train_docs=20000, seed=1005):
    """Serving workload: a training corpus of 20,000 docs (lengths LogNormal
    with mean 300, V=100,000, Zipf(1.07) background mixed into 500 Dir(0.01)
    topics) for a K=500 model, and `n_requests` requests of 1..64 unseen
    documents each, lengths LogNormal(ln 150, 1.0) clipped to [5, 10000],
    drawn from the same topics.  Unseen words get the training dictionary's
    ids; words the training corpus never saw are dropped (make_doc semantics).
    Returns (train_words, train_off, n_words), [(words, off) per request]."""
    V, true_k = 100000, 500
    rng = _rng(seed)
    topics, zcdf = _topics(rng, V, true_k, 0.01, 1.07)
    tl = lognormal_lengths(_rng(seed + 7919), train_docs, np.log(300.0) - 0.125, 0.5, 10, 3000)
    tw, toff = _draw(rng, topics, zcdf, V, true_k, 0.1, tl)
    dense, mapping = _first_appearance_ids(tw, V)
    rq = _rng(seed + 104729)
    sizes = rq.integers(1, 65, size=n_requests)
    lengths = lognormal_lengths(rq, int(sizes.sum()), np.log(150.0), 1.0, 5, 10000)
    uw, uoff = _draw(rng, topics, zcdf, V, true_k, 0.1, lengths)
    m = mapping[uw]
    keep = m >= 0
    nd = len(lengths)
    doc_id = np.repeat(np.arange(nd), np.diff(uoff).astype(np.int64))
    new_len = np.bincount(doc_id[keep], minlength=nd)
    uoff2 = np.zeros(nd + 1, dtype=np.uint64)
    np.cumsum(new_len, out=uoff2[1:])
    kept = m[keep].astype(np.uint32)
    reqs = []
    d0 = 0
    for n in sizes:
        d1 = d0 + int(n)
        b, e = int(uoff2[d0]), int(uoff2[d1])
        reqs.append((kept[b:e], (uoff2[d0:d1 + 1] - uoff2[d0]).astype(np.uint64)))
        d0 = d1
    return (dense.astype(np.uint32), toff, int((mapping >= 0).sum())), reqs


WORKLOADS = {
    "plain_k20": dict(seed=1001, n_docs=1000, V=5000, true_k=20, alpha=0.1, beta=0.01, mu=np.log(200.0), sigma=0.5,
                      lo=20, hi=2000, zipf_s=None, model_k=20, model_alpha=0.1, model_eta=0.01, iterations=100),
    "readme_k100": dict(seed=1002, n_docs=1000, V=50000, true_k=100, alpha=0.1, beta=0.01,
                        mu=np.log(1500.0) - 0.125, sigma=0.5, lo=10, hi=15000, zipf_s=1.07, model_k=100,
                        model_alpha=0.1, model_eta=0.01, iterations=200),
    "large_k1000": dict(seed=1003, n_docs=100000, V=100000, true_k=1000, alpha=0.05, beta=0.01,
                        mu=np.log(300.0) - 0.125, sigma=0.5, lo=10, hi=3000, zipf_s=1.07, model_k=1000,
                        model_alpha=0.05, model_eta=0.01, iterations=50),
}


def workload(name, n_docs=None):
    """Generate one of the named training corpora (optionally fewer docs)."""
    w = dict(WORKLOADS[name])
    D = int(n_docs or w["n_docs"])
    rng = _rng(w["seed"] + 7919)
    lengths = lognormal_lengths(rng, D, w["mu"], w["sigma"], w["lo"], w["hi"])
    words, off, nw = generate(w["seed"], D, w["V"], w["true_k"], w["alpha"], w["beta"], lengths, w["zipf_s"])
    return words, off, nw, w


def unseen_docs(seed, n_docs, V, true_k, mu, sigma, lo, hi, zipf_s=1.07, alpha=0.1, beta=0.01):
    """Unseen documents over word ids 0..V-1 (ids of an existing dictionary)."""
    rng = _rng(seed + 104729)
    lengths = lognormal_lengths(rng, n_docs, mu, sigma, lo, hi)
    words, off, nw = generate(seed, n_docs, V, true_k, alpha, beta, lengths, zipf_s)
    return words, off
