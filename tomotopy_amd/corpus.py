"""Host-side corpus preparation: vocabulary ordering and the token CSR.

Restates tomoto's prepare() data steps for plain LDA:
  * removeStopwords (src/TopicModel/TopicModel.hpp:526-571): words are ordered
    by (collection frequency, document frequency) descending with a stable
    sort; words below min_cf / min_df go after the kept ones; the rm_top most
    frequent kept words are rotated to the very end; realV is the number of
    kept words, and ids >= realV are skipped by the sampler.
  * sortAndWriteOrder (src/Utils/Utils.hpp:227-248): each document's words are
    stably sorted ascending and w_order[old_position] = new_position.
  * updateWeakArray (TopicModel.hpp:482-494): all documents' tokens in one
    array with per-document offsets (the CSR uploaded to the device).
"""
import numpy as np


def vocab_order(cf, df, min_cf=0, min_df=0, rm_top=0):
    """Return (order, real_v): order[old_id] = new_id, as removeStopwords."""
    cf = np.asarray(cf, dtype=np.int64)
    df = np.asarray(df, dtype=np.int64)
    n = len(cf)
    failed = (cf < min_cf) | (df < min_df)
    # stable sort: kept words first, each group by (cf, df) descending
    idx = np.lexsort((np.arange(n), -df, -cf, failed.astype(np.int64)))
    rot = min(int(rm_top), n)
    if rot:
        idx = np.concatenate([idx[rot:], idx[:rot]])
    sorted_failed = failed[idx]
    limit = n - rot
    hits = np.nonzero(sorted_failed[:limit])[0]
    real_v = int(hits[0]) if len(hits) else limit
    order = np.empty(n, dtype=np.int64)
    order[idx] = np.arange(n)
    return order, real_v


def counts(docs_words, n_words):
    """Collection and document frequency of every word id."""
    cf = np.zeros(n_words, dtype=np.int64)
    df = np.zeros(n_words, dtype=np.int64)
    for w in docs_words:
        w = np.asarray(w, dtype=np.int64)
        if len(w):
            cf += np.bincount(w, minlength=n_words)
            df[np.unique(w)] += 1
    return cf, df


def build_csr(docs_words, order):
    """Remap ids by `order`, sort each document, and concatenate.

    Returns (words uint32, doc_off uint64, w_order list of int arrays).
    """
    lens = np.fromiter((len(w) for w in docs_words), dtype=np.int64, count=len(docs_words))
    off = np.zeros(len(docs_words) + 1, dtype=np.uint64)
    np.cumsum(lens, out=off[1:])
    words = np.empty(int(off[-1]), dtype=np.uint32)
    w_orders = []
    for d, w in enumerate(docs_words):
        mapped = order[np.asarray(w, dtype=np.int64)] if len(w) else np.zeros(0, np.int64)
        perm = np.argsort(mapped, kind="stable")
        b = int(off[d])
        words[b:b + len(w)] = mapped[perm]
        wo = np.empty(len(w), dtype=np.int64)
        wo[perm] = np.arange(len(w))
        w_orders.append(wo)
    return words, off, w_orders


def csr_from_arrays(words, doc_off, order):
    """Vectorised build_csr for corpora already in CSR form (raw ids)."""
    words = np.asarray(words)
    doc_off = np.asarray(doc_off, dtype=np.uint64)
    D = len(doc_off) - 1
    mapped = np.asarray(order, dtype=np.int64)[words.astype(np.int64)]
    doc_id = np.repeat(np.arange(D, dtype=np.int64), np.diff(doc_off).astype(np.int64))
    perm = np.lexsort((np.arange(len(mapped)), mapped, doc_id))
    return mapped[perm].astype(np.uint32), doc_off.copy()


def prepare_arrays(words, doc_off, n_words, min_cf=0, min_df=0, rm_top=0):
    """tomoto prepare() on a raw-id CSR: returns (words, doc_off, real_v, order)."""
    words = np.asarray(words, dtype=np.int64)
    doc_off = np.asarray(doc_off, dtype=np.uint64)
    D = len(doc_off) - 1
    cf = np.bincount(words, minlength=n_words).astype(np.int64)
    doc_id = np.repeat(np.arange(D, dtype=np.int64), np.diff(doc_off).astype(np.int64))
    pairs = np.unique(doc_id * np.int64(n_words) + words)
    df = np.bincount(pairs % n_words, minlength=n_words).astype(np.int64)
    order, real_v = vocab_order(cf, df, min_cf, min_df, rm_top)
    w, off = csr_from_arrays(words, doc_off, order)
    return w, off, real_v, order


def term_weights(words, doc_off, real_v, scheme, vocab_w=None):
    """Per-token weights of TermWeight IDF or PMI, as tomoto's prepare()
    computes them (LDAModel.hpp:1096-1121 and 800-809).  vocab_w: the
    vocabulary weights to use (a trained model's, for unseen documents as
    initializeDocState does, LDAModel.hpp:794-809); None computes them from
    this corpus.

    words / doc_off: the prepared CSR (sorted documents, ids >= real_v are
    removed words).  Removed words get weight 0 but count in the document
    length and in the total collection frequency.
      IDF: w = float32(log(D / df[v])), df = documents containing v.
      PMI: vw[v] = float32(cf[v] / totCf) with totCf the number of tokens,
           w = max(float32(log(float32(float32(tf_d[v]) / vw[v]) / float32(len_d))), 0),
           tf_d[v] = occurrences of v in document d; the quotients are
           float32, the logarithm is taken in double.
    """
    if scheme not in ("idf", "pmi"):
        raise ValueError("scheme must be 'idf' or 'pmi'")
    words = np.asarray(words, dtype=np.int64)
    doc_off = np.asarray(doc_off, dtype=np.uint64)
    D = len(doc_off) - 1
    N = len(words)
    lens = np.diff(doc_off.astype(np.int64))
    doc_id = np.repeat(np.arange(D, dtype=np.int64), lens)
    valid = words < real_v
    out = np.zeros(N, dtype=np.float32)
    vw = vocab_weights(words, doc_off, real_v, scheme) if vocab_w is None else np.asarray(vocab_w, np.float32)
    if N == 0:
        return out
    vw_idx = np.where(valid, words, 0)
    if scheme == "idf":
        out[valid] = vw[vw_idx[valid]]
        return out
    # tf: the length of the run of equal ids (documents are sorted) each token lies in
    key = doc_id * np.int64(real_v + 1) + np.where(valid, words, real_v)
    starts = np.ones(N, dtype=bool)
    starts[1:] = key[1:] != key[:-1]
    run_id = np.cumsum(starts) - 1
    run_len = np.bincount(run_id)
    tf = run_len[run_id].astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        q = (tf / vw[vw_idx]) / lens[doc_id].astype(np.float32)
        w = np.log(q.astype(np.float64)).astype(np.float32)
    w = np.maximum(w, np.float32(0.0))
    out[valid] = w[valid]
    return out


def vocab_weights(words, doc_off, real_v, scheme):
    """tomoto's vocabWeights (LDAModel.hpp:1096-1121), one float32 per used
    word: IDF float32(log(D / df[v])) (inf for a word in no document), PMI
    float32(cf[v] / totCf) with totCf the number of tokens, removed words
    included.  Empty for scheme 'one'."""
    if scheme == "one":
        return np.zeros(0, np.float32)
    words = np.asarray(words, dtype=np.int64)
    doc_off = np.asarray(doc_off, dtype=np.uint64)
    D = len(doc_off) - 1
    N = len(words)
    valid = words < real_v
    if scheme == "idf":
        doc_id = np.repeat(np.arange(D, dtype=np.int64), np.diff(doc_off.astype(np.int64)))
        pairs = np.unique(doc_id[valid] * np.int64(real_v) + words[valid])
        df = np.bincount(pairs % max(real_v, 1), minlength=real_v).astype(np.float64)
        with np.errstate(divide="ignore"):
            return np.log(np.float64(D) / df).astype(np.float32)
    if scheme != "pmi":
        raise ValueError("scheme must be 'one', 'idf' or 'pmi'")
    cf = np.bincount(words[valid], minlength=real_v).astype(np.float64)
    return (cf / np.float64(max(N, 1))).astype(np.float32)
