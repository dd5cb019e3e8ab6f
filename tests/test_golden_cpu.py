"""The committed oracle trajectories (tests/golden/trajectories.json) are what
the oracle computes: one seed of the K=20 case regenerated in full, and the
first checkpoint of one K=100 seed, must equal the stored values exactly
(the partition schedule is deterministic for a seed and worker count,
whatever the number of threads)."""
import json
import os

import numpy as np

from oracle import oracle as O
from tomotopy_amd import corpus, synth

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "trajectories.json")


def _regen(case, seed, checkpoints, threads):
    words, off, nw, _ = synth.workload(case["workload"], n_docs=case["n_docs"])
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw)
    N = int(np.count_nonzero(ws < V))
    assert (V, len(offs) - 1, N) == (case["V"], case["D"], case["N"])
    m = O.Model(case["K"], V, case["alpha"], case["eta"], seed, scan_order=0)
    m.set_optim(case["optim_interval"], 0)
    m.set_corpus(ws, offs)
    out = []
    for _ in range(checkpoints):
        m.train(case["every"], workers=case["workers"], threads=threads)
        out.append(sum(m.loglik(0)) / N)
    return out


def test_k20_fixture_regenerates_exactly():
    with open(FIXTURE) as f:
        case = json.load(f)["plain_k20"]
    assert len(case["seeds"]) == 60
    got = _regen(case, 7, case["iterations"] // case["every"], threads=3)
    assert got == case["trajectories"]["7"]


def test_k100_fixture_first_checkpoint():
    with open(FIXTURE) as f:
        case = json.load(f)["readme_k100"]
    got = _regen(case, 2, 1, threads=4)
    assert got == case["trajectories"]["2"][:1]
