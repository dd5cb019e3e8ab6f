"""ll_per_word trajectories of GPU training against the CPU oracle's
partition schedule (tomoto's default scheme for LDA, LDAModel.hpp:166,
373-399), whose trajectories are committed in tests/golden/trajectories.json
(written by tests/golden/make_trajectories.py; the oracle is deterministic
for a corpus, seed and worker count).  Acceptance: the seed-averaged
ll_per_word within 1% of the oracle's at every checkpoint.

  plain_k20        1,000 docs, K=20, 100 iterations, alpha optimised every
                   10, 60 seeds per side.  Single seeds settle in optima
                   about 2% apart; with 60 seeds the standard error of the
                   difference of the means is about 0.12%.
  large_k1000_10k  the K=1000 corpus cut to 10,000 documents (3M tokens),
                   alpha 0.05, eta 0.01, 50 iterations, alpha every 10,
                   5 seeds per side: the device-wide schedule (at most an
                   eighth of the documents in flight) and tomoto's partition
                   schedule with the device's worker count.
  readme_k100      1,000 docs of mean length 1,500, V=50,000, K=100, 200
                   iterations, 10 oracle seeds: the device-wide schedule
                   over 10 seeds and the partition schedule with the
                   device's worker count over 3 (a partition sweep is bound
                   by the vocabulary chunk of the most frequent word, whose
                   tokens one worker samples in sequence: 60 ms here).
                   The chains are still descending steeply between
                   iterations 60 and 160, where a shift of a few iterations
                   moves ll_per_word by 1%; the check is at iteration 200,
                   the earlier checkpoints are printed.
"""
import json
import os

import numpy as np
import pytest

from tomotopy_amd import _native, corpus, synth

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "trajectories.json")


def _case(name):
    with open(FIXTURE) as f:
        case = json.load(f)[name]
    words, off, nw, _ = synth.workload(case["workload"], n_docs=case["n_docs"])
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw)
    assert (V, len(offs) - 1, int(np.count_nonzero(ws < V))) == (case["V"], case["D"], case["N"])
    return case, ws, offs, V


def _device_trajectories(case, ws, offs, V, workers, scheme, seeds=None):
    N = case["N"]
    out = []
    for seed in seeds or case["seeds"]:
        ctx = _native.Context(case["K"], V, case["alpha"], case["eta"], seed=seed)
        ctx.upload_corpus(ws, offs)
        ctx.set_optim(case["optim_interval"], 0)
        tr = []
        for _ in range(case["iterations"] // case["every"]):
            ctx.train(case["every"], workers=workers, scheme=scheme)
            tr.append(sum(ctx.loglik()) / N)
        assert not ctx.check_invariants().any()
        out.append(tr)
        ctx.close()
    return np.array(out)


def _compare(case, dev, label, checkpoints=None):
    cpu = np.array([case["trajectories"][str(s)] for s in case["seeds"]])
    g, c = dev.mean(axis=0), cpu.mean(axis=0)
    rel = (g - c) / np.abs(c)
    se = np.sqrt(dev.var(axis=0, ddof=1) / len(dev) + cpu.var(axis=0, ddof=1) / len(cpu)) / np.abs(c)
    print("%s: device - oracle, %% of ll_per_word per checkpoint: %s" % (label, np.round(100 * rel, 3).tolist()))
    print("%s: standard error of that difference, %%: %s" % (label, np.round(100 * se, 3).tolist()), flush=True)
    sel = rel if checkpoints is None else rel[checkpoints]
    assert np.all(np.abs(sel) < 0.01), rel


def test_k20_device_wide_60_seeds():
    case, ws, offs, V = _case("plain_k20")
    dev = _device_trajectories(case, ws, offs, V, workers=0, scheme=0)
    _compare(case, dev, "plain_k20 device-wide")


@pytest.mark.parametrize("scheme", [0, 3], ids=["device-wide", "partition"])
def test_k1000_10k_docs(scheme):
    case, ws, offs, V = _case("large_k1000_10k")
    ctx = _native.Context(case["K"], V, case["alpha"], case["eta"], seed=1)
    ctx.upload_corpus(ws, offs)
    print("schedule", ctx.resolve_schedule(0, scheme), flush=True)
    ctx.close()
    dev = _device_trajectories(case, ws, offs, V, workers=0, scheme=scheme)
    _compare(case, dev, "large_k1000_10k scheme %d" % scheme)


@pytest.mark.parametrize("scheme", [3, 0], ids=["partition", "device-wide"])
def test_readme_k100_200_iterations(scheme):
    case, ws, offs, V = _case("readme_k100")
    seeds = case["seeds"][:3] if scheme == 3 else None
    dev = _device_trajectories(case, ws, offs, V, workers=0, scheme=scheme, seeds=seeds)
    _compare(case, dev, "readme_k100 scheme %d" % scheme, checkpoints=[-1])
