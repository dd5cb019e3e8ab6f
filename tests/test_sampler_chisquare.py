"""The sampler's distribution (getZLikelihoods + prefixSum +
sampleFromDiscreteAcc, LDAModel.hpp:273-285, sample.hpp:61-189) against the
exact conditional of collapsed Gibbs sampling: one token drawn many times
with uniform u, its topic frequencies tested by chi-square against
p_k ∝ (n_dk + α_k)(n_kw + η)/(n_k + Vη) after the token's removal, for both
scan orders.  The same test rejects the conditional computed without the
removal (a plausible mistake), which shows it has the power to see one."""
import numpy as np
import pytest
from scipy.stats import chisquare

from oracle import oracle as O


def exact_conditional(ndk, nkw, nk, alpha, eta, V, zo, remove=True):
    a = ndk.astype(np.float64).copy()
    c = nkw.astype(np.float64).copy()
    e = nk.astype(np.float64).copy()
    if remove:
        a[zo] -= 1
        c[zo] -= 1
        e[zo] -= 1
    p = (a + alpha) * (c + eta) / (e + V * eta)
    return p / p.sum()


def chi2_pvalue(z, p, K):
    """Chi-square p-value of topic draws z against probabilities p, bins with
    an expected count below 5 merged into one."""
    n = len(z)
    obs = np.bincount(z.astype(np.int64), minlength=K).astype(np.float64)
    exp = n * p
    small = exp < 5
    if small.any():
        obs = np.append(obs[~small], obs[small].sum())
        exp = np.append(exp[~small], exp[small].sum())
        if exp[-1] == 0:
            obs, exp = obs[:-1], exp[:-1]
    return chisquare(obs, exp * (obs.sum() / exp.sum())).pvalue


def frozen_case(K, seed):
    """One document row, one word column and n_k for a vocabulary of V words,
    with the token's own topic zo holding a count of 1 in the document and the
    column, so that removing it changes the conditional strongly."""
    rng = np.random.default_rng(seed)
    V = 500
    alpha = (rng.random(K) * 0.3 + 0.02).astype(np.float32)
    eta = np.float32(0.01)
    ndk = rng.integers(0, 6, K).astype(np.int32)
    nkw = rng.integers(0, 4, K).astype(np.int32)
    zo = int(rng.integers(0, K))
    ndk[zo] = 1
    nkw[zo] = 1
    nk = (nkw + rng.integers(5, 40, K)).astype(np.int32)
    return V, alpha, eta, ndk, nkw, nk, zo


@pytest.mark.parametrize("K", [7, 20, 32])
@pytest.mark.parametrize("scan_order", [0, 1])
def test_oracle_sampler_matches_exact_conditional(K, scan_order):
    V, alpha, eta, ndk, nkw, nk, zo = frozen_case(K, seed=K)
    n = 200_000
    rng = np.random.default_rng(100 + K)
    u = (rng.integers(0, 1 << 24, n) * 2.0**-24).astype(np.float32)
    veta = np.float32(eta * np.float32(V))
    z, _ = O.sample_frozen(K, scan_order, alpha, eta, veta, ndk[None, :], nkw[None, :], nk,
                           np.zeros(n, np.int64), np.zeros(n, np.uint32), np.full(n, zo, np.uint16), u)
    p = exact_conditional(ndk, nkw, nk, alpha.astype(np.float64), float(eta), V, zo)
    assert chi2_pvalue(z, p, K) > 1e-3
    wrong = exact_conditional(ndk, nkw, nk, alpha.astype(np.float64), float(eta), V, zo, remove=False)
    assert chi2_pvalue(z, wrong, K) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("K", [7, 20, 32])
def test_device_sampler_matches_exact_conditional(K):
    """The device sampler (lda_sample_frozen) on a token whose own topic holds
    a count of 1 in its word column (and the least count in its document
    among such tokens), so that its removal changes the conditional strongly."""
    from tomotopy_amd import _native, corpus, synth
    words, off, nw, _ = synth.workload("plain_k20", n_docs=60)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw)
    rng = np.random.default_rng(K)
    alpha = (rng.random(K) * 0.3 + 0.02).astype(np.float32)
    eta = np.float32(0.01)
    z0 = rng.integers(0, K, len(ws)).astype(np.uint16)
    ctx = _native.Context(K, V, alpha, eta, seed=5)
    ctx.upload_corpus(ws, offs, z0)
    _, ndk, nkw, nk = ctx.download_state(z=False)
    doc_of = np.repeat(np.arange(len(offs) - 1), np.diff(offs).astype(np.int64))
    cand = [i for i in range(len(ws)) if ws[i] < V and nkw[ws[i], z0[i]] == 1]
    t = min(cand, key=lambda i: ndk[doc_of[i], z0[i]])
    d, v, zo = int(doc_of[t]), int(ws[t]), int(z0[t])
    n = 200_000
    u = (rng.integers(0, 1 << 24, n) * 2.0**-24).astype(np.float32)
    print("K", K, "token", t, "n", n, flush=True)
    z, _ = ctx.sample_frozen(np.full(n, d, np.int64), np.full(n, v, np.uint32), np.full(n, zo, np.uint16), u)
    p = exact_conditional(ndk[d], nkw[v], nk, alpha.astype(np.float64), float(eta), V, zo)
    assert chi2_pvalue(z, p, K) > 1e-3
    wrong = exact_conditional(ndk[d], nkw[v], nk, alpha.astype(np.float64), float(eta), V, zo, remove=False)
    assert chi2_pvalue(z, wrong, K) < 1e-6
