"""Autotuner of the batched fold-in launch: tokens per wavefront.

Batched fold-in groups documents into buckets by the number of wavefronts
that sample them, S = 1, 2, 4, 8, with bucket boundaries at seg_tokens,
2*seg_tokens and 4*seg_tokens tokens (the segmentation rule of the fold-in
kernel).  seg_tokens trades fidelity for speed: documents of at most
seg_tokens tokens are sampled sequentially, exactly as tomotopy does, while
longer ones are split into segments that see each other's n_dk changes one
iteration late.  The tuner times the batched launch for each candidate on a
sample of the workload, keeps the candidates whose log-likelihood per word
of the sample is within `ll_tol` (relative) of the exact sampler's
(seg_tokens = 0), and among those the largest seg_tokens (the most documents
sampled exactly) whose time is within `slack` of the fastest of them.
"""
import numpy as np

CANDIDATES = (0, 4096, 2048, 1024, 512, 256, 128, 64)


def _sample(words, doc_off, max_docs, seed, weights=None):
    """All of the longest documents plus a uniform sample of the rest (and
    their token weights, when given)."""
    D = len(doc_off) - 1
    if D <= max_docs:
        return words, doc_off, weights
    lens = np.diff(doc_off.astype(np.int64))
    n_long = max_docs // 4
    longest = np.argsort(-lens, kind="stable")[:n_long]
    rest = np.setdiff1d(np.arange(D), longest)
    rng = np.random.default_rng(seed)
    pick = np.sort(np.concatenate([longest, rng.choice(rest, size=max_docs - n_long, replace=False)]))
    sub_off = np.zeros(len(pick) + 1, np.uint64)
    sub_off[1:] = np.cumsum(lens[pick])
    sub_w = np.concatenate([words[int(doc_off[d]):int(doc_off[d + 1])] for d in pick]).astype(np.uint32)
    sub_t = None
    if weights is not None:
        sub_t = np.concatenate([weights[int(doc_off[d]):int(doc_off[d + 1])] for d in pick]).astype(np.float32)
    return sub_w, sub_off, sub_t


def tune_infer(ctx, words, doc_off, iterations, infer_seed=0, candidates=CANDIDATES, slack=0.10, max_docs=4000,
               repeats=2, seed=0, weights=None, ll_tol=1e-3):
    """Time the batched fold-in of a sample of (words, doc_off) for each
    candidate seg_tokens (with the largest segment count the kernel fits at
    this K) and set the chosen one on the context.  weights: the tokens'
    weights of a term-weighted model.  Candidates whose sample ll per word
    differs from seg_tokens = 0's by more than ll_tol (relative) are not
    chosen (0 must be among the candidates for the guard to apply).  Returns
    a dict with the choice and the per-candidate kernel times (ms) and
    log-likelihoods per word (per unit of weight) of the sample."""
    words = np.ascontiguousarray(words, np.uint32)
    doc_off = np.ascontiguousarray(doc_off, np.uint64)
    sw, so, st = _sample(words, doc_off, max_docs, seed, weights)
    n_valid = max(1, int(np.count_nonzero(sw < ctx.v)))
    if st is not None:
        n_valid = max(1e-30, float(np.sum(st[sw < ctx.v].astype(np.float64))))
    seg_max = ctx.infer_options()["seg_supported"]
    results = []
    for T in candidates:
        ctx.set_infer_options(int(T), seg_max if T else 1)
        e8 = int(so[min(8, len(so) - 1)])
        ctx.infer(sw[:e8], so[:min(8, len(so) - 1) + 1], 1, weights=None if st is None else st[:e8])  # graph capture
        best = None
        ll = None
        for _ in range(repeats):
            _, ll_d, _ = ctx.infer(sw, so, iterations, infer_seed=infer_seed, mode=0, weights=st)
            ms = ctx.kernel_stats()["ms"]
            best = ms if best is None else min(best, ms)
            ll = float(ll_d.sum()) / n_valid
        results.append({"seg_tokens": int(T), "seg_max": seg_max if T else 1, "kernel_ms": best, "ll_per_word": ll})
    exact = [r["ll_per_word"] for r in results if r["seg_tokens"] == 0]
    for r in results:
        r["ll_rel_diff"] = abs(r["ll_per_word"] - exact[0]) / abs(exact[0]) if exact else 0.0
    faithful = [r for r in results if r["ll_rel_diff"] <= ll_tol] or results
    fastest = min(r["kernel_ms"] for r in faithful)
    ok = [r for r in faithful if r["kernel_ms"] <= (1.0 + slack) * fastest]
    # largest seg_tokens among the acceptable ones; 0 (never segment) counts as the largest
    choice = max(ok, key=lambda r: (r["seg_tokens"] == 0, r["seg_tokens"]))
    ctx.set_infer_options(choice["seg_tokens"], choice["seg_max"])
    return {"seg_tokens": choice["seg_tokens"], "seg_max": choice["seg_max"], "slack": slack, "ll_tol": ll_tol,
            "sample_docs": int(len(so) - 1), "candidates": results}
