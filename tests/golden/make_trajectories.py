"""Writes tests/golden/trajectories.json: ll_per_word trajectories of the CPU
oracle's partition schedule (tomoto's default scheme for LDA,
LDAModel.hpp:166, 373-399) on the benchmark corpora, for the GPU trajectory
tests to compare against.

The oracle is deterministic for a given corpus, seed and worker count, so its
trajectories are fixed data; computing them takes minutes of CPU time, which
the GPU tests should not spend.  Each case records its corpus, sizes,
hyperparameters, the oracle worker count, and per seed the FP64 ll_per_word
(getLLDocs + getLLRest over the token count, LDAModel.hpp:616-674) after
every checkpoint.

    python tests/golden/make_trajectories.py [case ...]

Cases:
  plain_k20     1,000 docs, K=20, alpha 0.1, eta 0.01, 100 iterations,
                alpha optimised every 10, 60 seeds, 8 workers
  readme_k100   1,000 docs (about 1.5M tokens), V 50,000, K=100, alpha 0.1,
                eta 0.01, 200 iterations, alpha every 10, 10 seeds, 16 workers
  large_k1000_10k  the 100,000-document K=1000 corpus cut to its first
                10,000 documents (about 3M tokens), K=1000, alpha 0.05,
                eta 0.01, 50 iterations, alpha every 10, 5 seeds, 16 workers
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tomotopy_amd import corpus, synth  # noqa: E402

OUT = os.path.join(HERE, "trajectories.json")

CASES = {
    "plain_k20": dict(workload="plain_k20", n_docs=None, K=20, alpha=0.1, eta=0.01, iterations=100, every=10,
                      optim_interval=10, seeds=list(range(1, 61)), workers=8),
    "readme_k100": dict(workload="readme_k100", n_docs=None, K=100, alpha=0.1, eta=0.01, iterations=200, every=20,
                        optim_interval=10, seeds=list(range(1, 11)), workers=16),
    "large_k1000_10k": dict(workload="large_k1000", n_docs=10000, K=1000, alpha=0.05, eta=0.01, iterations=50,
                            every=10, optim_interval=10, seeds=[1, 2, 3, 4, 5], workers=16),
}


def corpus_of(case):
    words, off, nw, _ = synth.workload(case["workload"], n_docs=case["n_docs"])
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw)
    return ws, offs, V


def run_case(name, threads):
    case = CASES[name]
    ws, offs, V = corpus_of(case)
    N = int(np.count_nonzero(ws < V))
    out = dict(case)
    out.update(V=int(V), D=int(len(offs) - 1), N=N, scan_order="tomoto", trajectories={})
    for seed in case["seeds"]:
        t0 = time.time()
        m = O.Model(case["K"], V, case["alpha"], case["eta"], seed, scan_order=0)
        m.set_optim(case["optim_interval"], 0)
        m.set_corpus(ws, offs)
        traj = []
        for _ in range(case["iterations"] // case["every"]):
            m.train(case["every"], workers=case["workers"], threads=threads)
            traj.append(sum(m.loglik(0)) / N)
        out["trajectories"][str(seed)] = traj
        print(name, "seed", seed, "%.1fs" % (time.time() - t0), np.round(traj, 5).tolist(), flush=True)
    return out


def main():
    names = sys.argv[1:] or list(CASES)
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    threads = min(os.cpu_count() or 1, 16)
    for name in names:
        data[name] = run_case(name, threads)
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
