#!/usr/bin/env python3
"""Benchmark of the Gibbs sweep on one MI355X.

Workload: the large_k1000 training corpus (100,000 synthetic docs from the
LDA generative process with Zipf(1.07) background, mean length 300, about
30M tokens, V=100,000), model K=1000, alpha=0.05, eta=0.01.  One step is one
training iteration of LDAModel.train (one device-wide sweep over every token;
the alpha fixed point runs every 10th iteration as in tomotopy's default).

Prints one JSON line.  `value` is tokens sampled per second over the timed
steps (corpus and counts already resident in HBM).  `roofline` prices the
sweep kernel's algorithmic bytes per launch, N*(4K+8) + D*8K, against the
8 TB/s HBM peak, using the kernel's average duration from HIP events on the
library's stream.  `cpu_baseline` times the CPU oracle's partition sweep on
a bounded sample of the same corpus on this host's cores (`cpu_baseline_w1`:
one worker, tomotopy's `none` schedule).  `pmi_variant` repeats the training
measurement with TermWeight.PMI token weights.  The fold-in legs time the
oracle's fold-in on the host beside the device: the delta-state form
(`cpu_baseline`) and tomoto's own form with a K x V copy and a full
log-likelihood scan per document (`cpu_baseline_faithful`).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gibbs tokens/s at K=1000 (HBM %peak); batched infer docs/s + p99 vs plain"
HBM_PEAK_GBS = 8000.0


def bandwidth_probe(nbytes=1 << 30, reps=20):
    """The device's streaming ceilings from the library's hand-written probe
    kernels (lda_probe_bandwidth): a streaming read of 1 GiB in 16-byte
    vectors, and a copy (bytes read + bytes written), best of four grids."""
    from tomotopy_amd import _native
    rd, cp = _native.probe_bandwidth(nbytes, reps)
    return {"bytes": nbytes, "reps": reps, "read_GBps": rd, "copy_GBps": cp,
            "kernels": ["probe_read_kernel", "probe_copy_kernel"]}


def host_threads():
    """Worker threads for the CPU baselines: the host cores this process may
    use (OMP_NUM_THREADS where the host sets the share of its cores that one
    GPU's jobs get, else every core)."""
    n = os.cpu_count() or 1
    try:
        share = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        share = 0
    return max(1, min(n, share) if share > 0 else n)


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 byte
    counter summary (FETCH_SIZE and WRITE_SIZE runs of the bench's training
    leg, each scaled by what the counter reads for the probe copy of known
    size), with the commit the counters were taken on; (None, None) if the
    summary has no such kernel."""
    path = os.path.join(ROOT, "profiles", "traffic_large_k1000.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t["kernels"][kernel]["hbm_bytes_per_launch"], t.get("commit")
    except (OSError, KeyError, ValueError):
        return None, None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_corpus(n_docs=None, cache=None):
    from tomotopy_amd import corpus, synth
    t = time.time()
    if cache and os.path.exists(cache):
        z = np.load(cache)
        log("corpus loaded from %s" % cache)
        return z["words"], z["off"], int(z["V"]), synth.WORKLOADS["large_k1000"]
    words, off, nw, w = synth.workload("large_k1000", n_docs=n_docs)
    ws, offs, real_v, _ = corpus.prepare_arrays(words, off, nw)
    log("corpus: %d docs, %d tokens, V=%d (%.1fs)" % (len(offs) - 1, len(ws), real_v, time.time() - t))
    if cache:
        np.savez(cache, words=ws, off=offs, V=real_v)
    return ws, offs, real_v, w


def host_cpu_model():
    """The host CPU's model name, from /proc/cpuinfo."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ws, offs, V, K, alpha, eta, budget_s=12.0, threads=None, max_docs=4000, seed=11):
    """Oracle partition sweep with one worker per host thread (tomoto's
    default, W = the cores) over a random sample of max_docs documents
    (threads=1: tomotopy's `none` schedule on one core)."""
    from oracle import oracle as O
    threads = threads or host_threads()
    D = len(offs) - 1
    n_docs = max(1, min(D, max_docs))
    pick = np.sort(np.random.default_rng(seed).choice(D, size=n_docs, replace=False))
    lens = np.diff(offs.astype(np.int64))
    sub_off = np.zeros(n_docs + 1, np.uint64)
    sub_off[1:] = np.cumsum(lens[pick])
    sub_w = np.concatenate([ws[int(offs[d]):int(offs[d + 1])] for d in pick]).astype(np.uint32)
    m = O.Model(K, V, alpha, eta, seed=1, scan_order=0)
    m.set_optim(0, 0)
    m.set_corpus(sub_w, sub_off)
    m.train(1, workers=threads, threads=threads)  # warm (page in n_kw)
    sweeps = 0
    t0 = time.perf_counter()
    while True:
        m.train(1, workers=threads, threads=threads)
        sweeps += 1
        if time.perf_counter() - t0 > budget_s or sweeps >= 200:
            break
    dt = time.perf_counter() - t0
    ntok = int(np.count_nonzero(sub_w < V))
    return {
        "value": ntok * sweeps / dt,
        "unit": "tokens/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d random docs of %d (%d tokens, seed %d), K=%d, V=%d, %s scheme with %d workers, %d sweeps in "
                  "%.1fs" % (n_docs, D, ntok, seed, K, V, "partition" if threads > 1 else "none", threads, sweeps, dt),
        "cpu_model": host_cpu_model(), "host_cpus": os.cpu_count(),
    }


def narrow_hot(ws, V):
    """Words that keep int32 columns in the narrow-column sweep: every word id
    up to the last one with 2^15 or more tokens (the library's rule)."""
    cf = np.bincount(ws[ws < V].astype(np.int64), minlength=V)
    big = np.nonzero(cf >= 1 << 15)[0]
    return int(big[-1]) + 1 if len(big) else 0


def sweep_bytes(ws, offs, V, K, hot):
    """Algorithmic bytes of one device-wide sweep launch.  Per token the
    column read, 4K bytes for a word with an int32 column (ids < hot) and 2K
    for a word with a narrow u16 column, plus 8 (word id, topic read and
    written); per document 8K (n_dk staged in and written back).  `runs` is
    the same with one column read per run of equal words in a document (the
    sampler keeps the column while the word repeats)."""
    valid = ws < V
    v = ws.astype(np.int64)
    col = np.where(v < hot, 4 * K, 2 * K)
    D = len(offs) - 1
    N = int(np.count_nonzero(valid))
    per_doc = D * 8 * K
    tokens = int(col[valid].sum()) + 8 * N + per_doc
    start = np.ones(len(ws), dtype=bool)
    start[1:] = ws[1:] != ws[:-1]
    start[offs[:-1].astype(np.int64)[np.diff(offs.astype(np.int64)) > 0]] = True
    runs_mask = start & valid
    runs = int(col[runs_mask].sum()) + 8 * N + per_doc
    dense = N * (4 * K + 8) + per_doc
    return {"per_token": tokens, "per_run": runs, "dense_int32": dense, "runs": int(runs_mask.sum()), "tokens": N,
            "hot_words": hot}


def partition_chunks(ws, V, W):
    """Token counts of the W vocabulary chunks of tomoto's partition scheme
    (boundaries at cumulative collection-frequency quantiles,
    LDAModel.hpp:436-449, as the library computes them)."""
    cf = np.bincount(ws[ws < V].astype(np.int64), minlength=V)
    tot, cum, ch = int(cf.sum()), 0, []
    for i in range(V):
        cum += int(cf[i])
        if cum * W >= tot * (len(ch) + 1):
            ch.append(i + 1)
    while len(ch) < W:
        ch.append(V)
    ch = ch[:W]
    ch[-1] = V
    bounds = [0] + ch
    return [int(cf[bounds[j]:bounds[j + 1]].sum()) for j in range(W)]


def partition_bench(name, ws, offs, V, K, alpha, sweeps, cpu_sweeps=0):
    """tomoto's partition schedule on the device with its own worker count
    (train(workers=0, parallel=PARTITION)) beside the device-wide schedule on
    the same corpus.  One worker owns each vocabulary chunk for the whole
    sweep, so the chunk holding the most frequent words is sampled by one
    wavefront in sequence whatever W is: that chain bounds the sweep."""
    from tomotopy_amd import _native
    eta = 0.01
    N = int(np.count_nonzero(ws < V))
    D = len(offs) - 1
    ctx = _native.Context(K, V, alpha, eta, seed=1234)
    ctx.upload_corpus(ws, offs)
    ctx.set_optim(0, 0)
    _, W = ctx.resolve_schedule(0, 3)
    ctx.train(1, workers=0, scheme=3)  # graph capture
    t0 = time.perf_counter()
    ctx.train(sweeps, workers=0, scheme=3)
    dt = (time.perf_counter() - t0) / sweeps
    kms = ctx.kernel_stats()["ms"] / sweeps
    chunks = partition_chunks(ws, V, W)
    dw_sched = ctx.resolve_schedule(0, 0)
    ctx.train(1, workers=0, scheme=0)
    t0 = time.perf_counter()
    ctx.train(sweeps, workers=0, scheme=0)
    dt_dw = (time.perf_counter() - t0) / sweeps
    dense = N * (4 * K + 8) + D * 8 * K
    out = {"workload": name, "K": K, "V": V, "docs": D, "tokens": N, "workers": W, "rounds_per_sweep": W,
           "value": N / dt, "unit": "tokens/s", "ms_per_sweep": dt * 1e3, "kernel_ms_per_sweep": kms,
           "frac_dense_formula": dense / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "largest_chunk_tokens": max(chunks), "mean_chunk_tokens": N / W,
           "us_per_token_on_largest_chunk": kms * 1e3 / max(chunks),
           "device_wide": {"value": N / dt_dw, "ms_per_sweep": dt_dw * 1e3, "workgroups": dw_sched[1]}}
    ctx.close()
    if cpu_sweeps:
        from oracle import oracle as O
        thr = host_threads()
        m = O.Model(K, V, alpha, eta, seed=1, scan_order=0)
        m.set_optim(0, 0)
        m.set_corpus(ws, offs)
        m.train(1, workers=thr, threads=thr)
        t0 = time.perf_counter()
        m.train(cpu_sweeps, workers=thr, threads=thr)
        dtc = (time.perf_counter() - t0) / cpu_sweeps
        out["cpu_baseline"] = {"value": N / dtc, "unit": "tokens/s", "cores": thr, "kind": "port",
                               "sample": "the whole corpus, partition scheme with %d workers, %d sweeps" % (thr, cpu_sweeps),
                               "cpu_model": host_cpu_model()}
    log("partition %s: W=%d, %.3g tokens/s (%.1f ms/sweep; largest chunk %d tokens = %.2f us/token), device-wide "
        "%.3g tokens/s" % (name, W, out["value"], dt * 1e3, max(chunks), out["us_per_token_on_largest_chunk"],
                            out["device_wide"]["value"]))
    return out


def pmi_bench(ws, offs, V, K, alpha, eta, steps, warmup):
    """The same corpus and model with TermWeight.PMI token weights: float
    counts, clamped removals and the weighted merge (n_k rebuilt from n_kw)
    after every sweep.  Algorithmic bytes per sweep launch: one float column
    read (4K) per run of equal words in a document, 12 per token (word, topic
    read and written, weight), 8K per document; `frac_per_token` counts a
    column read for every token instead (N*(4K+12) + D*8K), which exceeds what
    reaches HBM since a run's tokens share one read."""
    from tomotopy_amd import _native, corpus
    t = time.time()
    wts = corpus.term_weights(ws, offs, V, "pmi")
    D = len(offs) - 1
    n_valid = int(np.count_nonzero(ws < V))
    ctx = _native.Context(K, V, alpha, eta, seed=1234, tw=2)
    ctx.upload_corpus(ws, offs, weights=wts)
    log("PMI weights + upload: %.2fs" % (time.time() - t))
    for i in range(warmup):
        ctx.train(1, workers=0)
        log("PMI warmup step %d: %.2f ms kernel" % (i, ctx.kernel_stats()["ms"]))
    t0 = time.perf_counter()
    ctx.train(steps, workers=0)
    dt = time.perf_counter() - t0
    st = ctx.kernel_stats()
    kern_ms = st["ms"] / max(1, st["launches"])
    per_token_bytes = n_valid * (4 * K + 12) + D * 8 * K
    sb = sweep_bytes(ws, offs, V, K, V)  # every column float: 4K per run
    bytes_per_launch = sb["per_run"] + 4 * n_valid
    achieved = bytes_per_launch / (kern_ms * 1e-3) / 1e9
    lld, llr = ctx.loglik()
    wn = ctx.weighted_n()
    out = {"value": n_valid * steps / dt, "unit": "tokens/s", "ms_per_step": dt * 1e3 / steps, "steps": steps,
           "kernel": "sweep_device_w_kernel<16, false>", "kernel_ms": kern_ms, "achieved_GBs": achieved,
           "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": bytes_per_launch, "runs": sb["runs"],
           "algorithmic_bytes_per_launch_per_token": per_token_bytes,
           "frac_per_token": per_token_bytes / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "traffic": pmc_traffic("sweep_device_w_kernel<16, false>")[0],
           "ll_per_word": (lld + llr) / wn, "weighted_n": wn}
    log("PMI: %.3g tokens/s, kernel %.3f ms (%.1f%% of HBM peak), %.3f ms/step, ll_per_word %.5f"
        % (out["value"], kern_ms, 100 * out["frac"], out["ms_per_step"], out["ll_per_word"]))
    return out


def cpu_infer_baseline(ctx, ws, offs, V, K, alpha, eta, uws, uoffs, iterations, budget_s=8.0, seed=3,
                       faithful=False):
    """The oracle's fold-in (tomoto's per-document infer, together=False) on
    the host's cores: documents drawn at random from the workload, split over
    threads (the oracle releases the GIL), until the time budget is spent.
    faithful=True runs tomoto's own form (LDAModel.hpp:931, 945): a private
    copy of the whole K x V state per document and the log-likelihood scan
    over all of it; otherwise the private state is a delta over the frozen
    model."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    threads = host_threads()
    z, _, _, _ = ctx.download_state(ndk=False, nkw=False, nk=False)
    m = O.Model(K, V, ctx.get_alpha(), eta, seed=1, scan_order=0)
    m.set_corpus(ws, offs, z)
    D = len(uoffs) - 1
    order = np.random.default_rng(seed).permutation(D)
    lens = np.diff(uoffs.astype(np.int64))

    def run(docs):
        sub_off = np.zeros(len(docs) + 1, np.uint64)
        sub_off[1:] = np.cumsum(lens[docs])
        sub_w = np.concatenate([uws[int(uoffs[d]):int(uoffs[d + 1])] for d in docs]).astype(np.uint32)
        if faithful:
            m.infer_gen(sub_w, sub_off, iterations, faithful=2)
        else:
            m.infer(sub_w, sub_off, iterations, mode=1)
        return len(docs), int(np.count_nonzero(sub_w < V))

    done = toks = 0
    pos = 0
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        while time.perf_counter() - t0 < budget_s and pos < D:
            # tomoto evaluates getLLRest(global) once per infer call; 16 documents per
            # call keep that scan a small share of the faithful form's work
            step, tasks = (16, threads) if faithful else (4, threads * 4)
            batch = [b for b in (order[pos + j:pos + j + step] for j in range(0, tasks * step, step)) if len(b)]
            pos += tasks * step
            for n, t in ex.map(run, batch):
                done += n
                toks += t
    dt = time.perf_counter() - t0
    form = ("tomoto's form: a private copy of the K x V state per document and the log-likelihood scan over "
            "all of it with the lgamma table, getLLRest(global) once per 16 documents" if faithful else
            "per-document private state as a delta over the frozen model; tomoto additionally copies "
            "the K x V state per document")
    return {"value": done / dt, "unit": "docs/s", "cores": threads, "kind": "port",
            "sample": "%d random docs (%d tokens) of the workload, %d iterations each, oracle fold-in "
                      "(%s), %d threads, %.1fs" % (done, toks, iterations, form, threads, dt),
            "cpu_model": host_cpu_model()}


def infer_bench(plain_sample=300, iterations=100):
    """Fold-in: batched (one persistent launch per segment bucket, replayed
    from a hipGraph) with the autotuned tokens per wavefront, against plain
    (one launch + synchronisation per document, same segmentation) and
    against the sequential sampler (no segmentation) for ll_per_word, on a
    frozen K=200 model."""
    from tomotopy_amd import _native, autotune, corpus, synth
    t = time.time()
    (tw, toff, nw), (uw, uoff) = synth.foldin_workload()
    ws, offs, V, order = corpus.prepare_arrays(tw, toff, nw)
    uws, uoffs = corpus.csr_from_arrays(uw, uoff, order)
    D = len(uoffs) - 1
    log("fold-in workload: %d unseen docs, %d tokens (%.1fs)" % (D, len(uws), time.time() - t))
    K = 200
    ctx = _native.Context(K, V, 0.1, 0.01, seed=77)
    ctx.upload_corpus(ws, offs)
    ctx.train(100, workers=0)
    ntok = int(np.count_nonzero(uws < V))
    # sequential sampler (every document exactly as tomotopy samples it)
    ctx.set_infer_options(0, 1)
    ctx.infer(uws[:int(uoffs[64])], uoffs[:65], 2)
    t0 = time.perf_counter()
    _, ll_seq, _ = ctx.infer(uws, uoffs, iterations, mode=0)
    seq_s = time.perf_counter() - t0
    tune = autotune.tune_infer(ctx, uws, uoffs, iterations)
    log("autotuner: seg_tokens %d seg_max %d (%s)" % (tune["seg_tokens"], tune["seg_max"], ", ".join(
        "%d:%.0fms" % (c["seg_tokens"], c["kernel_ms"]) for c in tune["candidates"])))
    ctx.infer(uws[:int(uoffs[64])], uoffs[:65], 2)
    t0 = time.perf_counter()
    theta, ll, _ = ctx.infer(uws, uoffs, iterations, mode=0)
    batched_s = time.perf_counter() - t0
    kst = ctx.kernel_stats()
    rng = np.random.default_rng(5)
    pick = np.sort(rng.choice(D, size=min(plain_sample, D), replace=False))
    lens = np.diff(uoffs.astype(np.int64))
    sub_off = np.zeros(len(pick) + 1, np.uint64)
    sub_off[1:] = np.cumsum(lens[pick])
    sub_w = np.concatenate([uws[int(uoffs[d]):int(uoffs[d + 1])] for d in pick]).astype(np.uint32)
    t0 = time.perf_counter()
    th_p, ll_p, _ = ctx.infer(sub_w, sub_off, iterations, mode=1)
    plain_s = time.perf_counter() - t0
    same = bool(np.array_equal(th_p, theta[pick]) and np.array_equal(ll_p, ll[pick]))
    llw, llw_seq = float(ll.sum() / ntok), float(ll_seq.sum() / ntok)
    out = {
        "workload": "foldin_k200", "docs": D, "tokens": ntok, "iterations": iterations, "K": K, "V": V,
        "seg_tokens": tune["seg_tokens"], "seg_max": tune["seg_max"], "autotune": tune,
        "batched_docs_per_s": D / batched_s, "batched_s": batched_s, "batched_kernel_ms": kst["ms"],
        "sequential_docs_per_s": D / seq_s, "sequential_s": seq_s,
        "plain_docs_per_s": len(pick) / plain_s, "plain_sample_docs": int(len(pick)),
        "speedup": (D / batched_s) / (len(pick) / plain_s),
        "ll_per_word": llw, "ll_per_word_sequential": llw_seq,
        "ll_per_word_rel_diff": abs(llw - llw_seq) / abs(llw_seq),
        "plain_equals_batched": same, "longest_doc": int(lens.max()),
    }
    out["cpu_baseline"] = cpu_infer_baseline(ctx, ws, offs, V, K, 0.1, 0.01, uws, uoffs, iterations)
    log("fold-in CPU baseline: %.0f docs/s (%s)" % (out["cpu_baseline"]["value"], out["cpu_baseline"]["sample"]))
    out["cpu_baseline_faithful"] = cpu_infer_baseline(ctx, ws, offs, V, K, 0.1, 0.01, uws, uoffs, iterations,
                                                      budget_s=6.0, faithful=True)
    log("fold-in CPU baseline, tomoto's form: %.1f docs/s (%s)"
        % (out["cpu_baseline_faithful"]["value"], out["cpu_baseline_faithful"]["sample"]))
    log("fold-in: batched %.0f docs/s (%.3fs), sequential %.0f docs/s, plain %.1f docs/s on %d sampled docs, "
        "identical=%s, ll_per_word %.5f vs sequential %.5f"
        % (out["batched_docs_per_s"], batched_s, out["sequential_docs_per_s"], out["plain_docs_per_s"], len(pick),
           same, llw, llw_seq))
    return out


def infer_pmi_bench(iterations=100, plain_sample=200, cpu_budget_s=8.0):
    """Fold-in of a TermWeight.PMI model (K=200, the fold-in workload): the
    token weights of the unseen documents from the model's vocabulary
    weights, each document on a private copy of the float counts, batched
    with the autotuned segmentation (and sequentially), against plain (one
    launch per document) and the oracle's fold-in on the host cores."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    from tomotopy_amd import _native, autotune, corpus, synth
    (tw, toff, nw), (uw, uoff) = synth.foldin_workload()
    ws, offs, V, order = corpus.prepare_arrays(tw, toff, nw)
    uws, uoffs = corpus.csr_from_arrays(uw, uoff, order)
    D = len(uoffs) - 1
    K = 200
    wts = corpus.term_weights(ws, offs, V, "pmi")
    vw = corpus.vocab_weights(ws, offs, V, "pmi")
    uwt = corpus.term_weights(uws, uoffs, V, "pmi", vocab_w=vw)
    ctx = _native.Context(K, V, 0.1, 0.01, seed=77, tw=2)
    ctx.upload_corpus(ws, offs, weights=wts)
    ctx.train(100, workers=0)
    ntok = int(np.count_nonzero(uws < V))
    wsum = float(np.sum(uwt[uws < V].astype(np.float64)))
    e64 = int(uoffs[64])
    ctx.infer(uws[:e64], uoffs[:65], 2, weights=uwt[:e64])
    # sequential sampler (every document exactly as tomotopy samples it)
    t0 = time.perf_counter()
    _, ll_seq, _ = ctx.infer(uws, uoffs, iterations, mode=0, weights=uwt)
    seq_s = time.perf_counter() - t0
    tune = autotune.tune_infer(ctx, uws, uoffs, iterations, weights=uwt)
    log("PMI autotuner: seg_tokens %d seg_max %d (%s)" % (tune["seg_tokens"], tune["seg_max"], ", ".join(
        "%d:%.0fms" % (c["seg_tokens"], c["kernel_ms"]) for c in tune["candidates"])))
    t0 = time.perf_counter()
    theta, ll, _ = ctx.infer(uws, uoffs, iterations, mode=0, weights=uwt)
    batched_s = time.perf_counter() - t0
    kst = ctx.kernel_stats()
    rng = np.random.default_rng(5)
    pick = np.sort(rng.choice(D, size=min(plain_sample, D), replace=False))
    lens = np.diff(uoffs.astype(np.int64))
    sub_off = np.zeros(len(pick) + 1, np.uint64)
    sub_off[1:] = np.cumsum(lens[pick])
    sub_w = np.concatenate([uws[int(uoffs[d]):int(uoffs[d + 1])] for d in pick]).astype(np.uint32)
    sub_t = np.concatenate([uwt[int(uoffs[d]):int(uoffs[d + 1])] for d in pick]).astype(np.float32)
    t0 = time.perf_counter()
    th_p, ll_p, _ = ctx.infer(sub_w, sub_off, iterations, mode=1, weights=sub_t)
    plain_s = time.perf_counter() - t0
    same = bool(np.array_equal(th_p, theta[pick]) and np.array_equal(ll_p, ll[pick]))
    llw, llw_seq = float(ll.sum() / wsum), float(ll_seq.sum() / wsum)
    out = {"workload": "foldin_k200_pmi", "docs": D, "tokens": ntok, "iterations": iterations, "K": K, "V": V,
           "seg_tokens": tune["seg_tokens"], "seg_max": tune["seg_max"],
           "batched_docs_per_s": D / batched_s, "batched_s": batched_s, "batched_kernel_ms": kst["ms"],
           "sequential_docs_per_s": D / seq_s, "sequential_s": seq_s,
           "plain_docs_per_s": len(pick) / plain_s, "plain_sample_docs": int(len(pick)),
           "speedup": (D / batched_s) / (len(pick) / plain_s), "plain_equals_batched": same,
           "ll_per_weight": llw, "ll_per_weight_sequential": llw_seq,
           "ll_per_weight_rel_diff": abs(llw - llw_seq) / abs(llw_seq)}
    # the oracle's fold-in on the host's cores, on the device's trained state
    threads = host_threads()
    z, ndk, nkw, nk = ctx.download_state_f()
    m = O.WeightedModel(K, V, ctx.get_alpha(), 0.01, seed=1, scan_order=0)
    m.set_corpus(ws, offs, wts, z)
    m.set_state(z, ndk, nkw, nk)
    order_d = rng.permutation(D)

    def run(docs):
        so = np.zeros(len(docs) + 1, np.uint64)
        so[1:] = np.cumsum(lens[docs])
        sw_ = np.concatenate([uws[int(uoffs[d]):int(uoffs[d + 1])] for d in docs]).astype(np.uint32)
        st_ = np.concatenate([uwt[int(uoffs[d]):int(uoffs[d + 1])] for d in docs]).astype(np.float32)
        m.infer_gen(sw_, so, iterations, weights=st_)
        return len(docs)

    done = pos = 0
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        while time.perf_counter() - t0 < cpu_budget_s and pos < D:
            batch = [b for b in (order_d[pos + j:pos + j + 4] for j in range(0, threads * 16, 4)) if len(b)]
            pos += threads * 16
            done += sum(ex.map(run, batch))
    dt = time.perf_counter() - t0
    out["cpu_baseline"] = {"value": done / dt, "unit": "docs/s", "cores": threads, "kind": "port",
                           "sample": "%d random docs of the workload, %d iterations each, oracle fold-in of the "
                                     "PMI model (private columns per document), %d threads, %.1fs"
                                     % (done, iterations, threads, dt)}
    log("PMI fold-in: batched %.0f docs/s (%.3fs, kernels %.1f ms), sequential %.0f docs/s, plain %.1f docs/s on %d "
        "docs, identical=%s, ll per weight %.5f vs sequential %.5f; CPU %.0f docs/s"
        % (out["batched_docs_per_s"], batched_s, kst["ms"], out["sequential_docs_per_s"], out["plain_docs_per_s"],
           len(pick), same, llw, llw_seq, out["cpu_baseline"]["value"]))
    return out


def _percentile(x, q):
    return float(np.percentile(np.asarray(x, dtype=np.float64), q)) if len(x) else None


def serve_setup(n_pool=600):
    """The serving workload (600 pooled requests) and its K=500, V=100,000
    model trained 50 iterations on the device."""
    from tomotopy_amd import _native, corpus, synth
    t = time.time()
    (tw, toff, nw), reqs_raw = synth.serve_workload(n_pool)
    ws, offs, V, order = corpus.prepare_arrays(tw, toff, nw)
    reqs = [corpus.csr_from_arrays(w, o, order) for w, o in reqs_raw]
    log("serving workload: %d training docs, %d pooled requests (%d docs) (%.1fs)"
        % (len(offs) - 1, len(reqs), sum(len(o) - 1 for _, o in reqs), time.time() - t))
    ctx = _native.Context(500, V, 0.1, 0.01, seed=55)
    ctx.upload_corpus(ws, offs)
    ctx.train(50, workers=0)
    allw = np.concatenate([w for w, _ in reqs]).astype(np.uint32)
    alloff = np.zeros(sum(len(o) - 1 for _, o in reqs) + 1, np.uint64)
    alloff[1:] = np.cumsum(np.concatenate([np.diff(o.astype(np.int64)) for _, o in reqs]))
    return ctx, ws, offs, V, reqs, allw, alloff


def serve_bench(setup=None, iterations=50, qps_list=(1, 2, 5, 10, 20, 50, 100, 200, 400, 800, 1200, 1600),
                seconds=4.0, plain_requests=60, seed=2024, seg_tokens=None, step_tokens=12000, lanes=1,
                lanes_sweep=(2, 4), ll_tol=1e-3, cpu=True):
    """Serving loop on a device-resident K=500, V=100,000 model: Poisson request
    arrivals at each QPS (requests of 1..64 docs, lengths LogNormal(ln 150, 1)
    clipped to [5, 10000], 50 iterations); latency is submit to results on the
    host, measured inside the server.  The segmentation is the autotuner's
    choice on the pooled requests' documents among the candidates whose ll
    per word stays within ll_tol of exact sampling (None: the fastest, no ll
    guard); the served ll per word is set against exact sampling (no
    segmentation) of the same requests.  Baselines on the same requests,
    served one at a time back to back: plain per-request GPU infer (one
    lda_infer call per request, with the served segmentation and exact), and
    (cpu=True) plain per-document GPU infer (one launch and one
    synchronisation per document) and the CPU oracle.  The top rate is
    repeated with each lane count of lanes_sweep."""
    import threading
    from tomotopy_amd import _native, autotune
    ctx, ws, offs, V, reqs, allw, alloff = setup or serve_setup()
    K = ctx.k
    tune = None
    if seg_tokens is None:
        tune = autotune.tune_infer(ctx, allw, alloff, iterations, infer_seed=seed,
                                   ll_tol=float("inf") if ll_tol is None else ll_tol)
        seg_tokens, seg_max = tune["seg_tokens"], tune["seg_max"]
        log("serve autotuner (ll guard %s): seg_tokens %d seg_max %d (%s)" % (ll_tol, seg_tokens, seg_max, ", ".join(
            "%d:%.0fms/%.2e" % (c["seg_tokens"], c["kernel_ms"], c["ll_rel_diff"]) for c in tune["candidates"])))
    else:
        seg_max = ctx.infer_options()["seg_supported"] if seg_tokens else 1
    ctx.set_infer_options(seg_tokens, seg_max)
    out = {"workload": "serve_k500", "K": K, "V": V, "iterations": iterations, "levels": [],
           "seg_tokens": seg_tokens, "seg_max": seg_max, "step_tokens": step_tokens, "lanes": lanes,
           "ll_guard": ll_tol, "autotune": tune}
    srv = _native.Server(ctx, hbm_budget_bytes=8 << 30, iterations=iterations, step_iters=iterations,
                         step_tokens=step_tokens, infer_seed=seed, lanes=lanes)

    def level(srv, qps, rng):
        n = max(8, min(len(reqs), int(qps * seconds)))  # at least 8 requests: 8 s at 1 request/s
        gaps = rng.exponential(1.0 / qps, size=n)
        st0 = srv.stats()
        pick = rng.integers(0, len(reqs), size=n)
        ids = [None] * n
        t_start = time.perf_counter()

        def client():
            due = t_start
            for j in range(n):
                due += gaps[j]
                dt = due - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                ids[j] = srv.submit(*reqs[pick[j]])

        th = threading.Thread(target=client)
        th.start()
        lat = []
        docs = 0
        lls, toks = 0.0, 0
        for j in range(n):
            while ids[j] is None:
                time.sleep(0.0005)
            _, ll, ms = srv.wait(ids[j])
            lat.append(ms)
            docs += len(ll)
            lls += float(ll.sum())
            toks += int(np.count_nonzero(reqs[pick[j]][0] < V))
        t_end = time.perf_counter()
        th.join()
        st1 = srv.stats()
        nst = max(1, st1["steps"] - st0["steps"])
        lvl = {"qps": qps, "requests": n, "docs": docs, "docs_per_s": docs / (t_end - t_start),
               "p50_ms": _percentile(lat, 50), "p99_ms": _percentile(lat, 99), "max_ms": max(lat),
               "steps": st1["steps"] - st0["steps"], "step_gpu_ms": (st1["gpu_ms"] - st0["gpu_ms"]) / nst,
               "step_host_ms": (st1["host_ms"] - st0["host_ms"]) / nst}
        log("serve qps %4d: %5d req, %7.0f docs/s, p50 %.1f ms, p99 %.1f ms; %d steps, %.2f ms device + %.2f ms host each"
            % (qps, n, lvl["docs_per_s"], lvl["p50_ms"], lvl["p99_ms"], lvl["steps"], lvl["step_gpu_ms"],
               lvl["step_host_ms"]))
        return lvl, lls, toks

    try:
        for i in range(4):  # warm
            srv.wait(srv.submit(*reqs[i]))
        rng = np.random.default_rng(seed)
        ll_sum, tok_sum = 0.0, 0
        for qps in qps_list:
            lvl, a, b = level(srv, qps, rng)
            out["levels"].append(lvl)
            ll_sum += a
            tok_sum += b
        # saturation: every pooled request submitted at once
        t0 = time.perf_counter()
        ids = [srv.submit(*r) for r in reqs]
        sat = [srv.wait(i) for i in ids]
        lat = [r[2] for r in sat]
        served_ll = [np.asarray(r[1]) for r in sat]
        dt = time.perf_counter() - t0
        ndocs = sum(len(o) - 1 for _, o in reqs)
        out["saturation"] = {"requests": len(reqs), "docs": ndocs, "docs_per_s": ndocs / dt,
                             "p50_ms": _percentile(lat, 50), "p99_ms": _percentile(lat, 99)}
        log("serve saturation: %.0f docs/s over %d requests" % (ndocs / dt, len(reqs)))
        st = srv.stats()
        out["server_steps"] = st["steps"]
        out["ll_per_word"] = ll_sum / max(tok_sum, 1)
    finally:
        srv.stop()
    out["lanes_sweep"] = []
    for ln in lanes_sweep:
        srv2 = _native.Server(ctx, hbm_budget_bytes=8 << 30, iterations=iterations, step_iters=iterations,
                              step_tokens=step_tokens, infer_seed=seed, lanes=ln)
        try:
            for i in range(4):
                srv2.wait(srv2.submit(*reqs[i]))
            lvl, _, _ = level(srv2, qps_list[-1], np.random.default_rng(seed + ln))
        finally:
            srv2.stop()
        lvl["lanes"] = ln
        out["lanes_sweep"].append(lvl)
    # served ll against exact sampling (no segmentation) of the same requests
    served_tok = int(np.count_nonzero(allw < V))
    out["served_ll_per_word"] = float(sum(float(np.sum(x)) for x in served_ll)) / max(served_tok, 1)
    if seg_tokens:
        ctx.set_infer_options(0, 1)
        _, ll_exact, _ = ctx.infer(allw, alloff, iterations, infer_seed=seed, mode=0)
        ctx.set_infer_options(seg_tokens, seg_max)
        out["exact_ll_per_word"] = float(ll_exact.sum()) / max(served_tok, 1)
    else:
        out["exact_ll_per_word"] = out["served_ll_per_word"]
    out["served_vs_exact_ll_rel_diff"] = abs(out["served_ll_per_word"] - out["exact_ll_per_word"]) / abs(
        out["exact_ll_per_word"])
    log("served ll per word %.5f vs exact sampling %.5f (rel %.2e)" % (
        out["served_ll_per_word"], out["exact_ll_per_word"], out["served_vs_exact_ll_rel_diff"]))
    # baselines: the same requests served one at a time
    sub = reqs[:plain_requests]
    lat_req = []
    t0 = time.perf_counter()
    ll_sum, tok_sum = 0.0, 0
    same_ll = True
    for j, (w, o) in enumerate(sub):
        t1 = time.perf_counter()
        _, ll, _ = ctx.infer(w, o, iterations, infer_seed=seed, mode=0)
        lat_req.append((time.perf_counter() - t1) * 1e3)
        ll_sum += float(ll.sum())
        tok_sum += int(np.count_nonzero(w < V))
        same_ll = same_ll and bool(np.array_equal(ll, served_ll[j]))
    dt = time.perf_counter() - t0
    ndocs = sum(len(o) - 1 for _, o in sub)
    out["plain_per_request"] = {"requests": len(sub), "docs": ndocs, "docs_per_s": ndocs / dt,
                                "p50_ms": _percentile(lat_req, 50), "p99_ms": _percentile(lat_req, 99),
                                "ll_per_word": ll_sum / max(tok_sum, 1), "seg_tokens": seg_tokens,
                                "served_ll_identical": same_ll}
    if seg_tokens:
        ctx.set_infer_options(0, 1)
        lat_x = []
        t0 = time.perf_counter()
        for w, o in sub:
            t1 = time.perf_counter()
            ctx.infer(w, o, iterations, infer_seed=seed, mode=0)
            lat_x.append((time.perf_counter() - t1) * 1e3)
        dtx = time.perf_counter() - t0
        ctx.set_infer_options(seg_tokens, seg_max)
        out["plain_per_request_exact"] = {"requests": len(sub), "docs": ndocs, "docs_per_s": ndocs / dtx,
                                          "p50_ms": _percentile(lat_x, 50), "p99_ms": _percentile(lat_x, 99),
                                          "seg_tokens": 0}
    if cpu:
        sub2 = reqs[:max(1, plain_requests // 10)]
        t0 = time.perf_counter()
        for w, o in sub2:
            ctx.infer(w, o, iterations, infer_seed=seed, mode=1)
        dt = time.perf_counter() - t0
        nd2 = sum(len(o) - 1 for _, o in sub2)
        out["plain_per_document"] = {"requests": len(sub2), "docs": nd2, "docs_per_s": nd2 / dt}
        out["cpu_baseline"] = cpu_infer_baseline(ctx, ws, offs, V, K, 0.1, 0.01, allw, alloff, iterations)
        log("serving CPU baseline: %.0f docs/s (%s)" % (out["cpu_baseline"]["value"], out["cpu_baseline"]["sample"]))
    best = max(lv["docs_per_s"] for lv in out["levels"])
    ok = [lv for lv in out["levels"] if lv["p99_ms"] is not None and lv["p99_ms"] <= 50.0]
    out["best_docs_per_s"] = max(best, out["saturation"]["docs_per_s"])
    out["speedup_vs_plain_per_request"] = out["best_docs_per_s"] / out["plain_per_request"]["docs_per_s"]
    out["max_qps_with_p99_le_50ms"] = max((lv["qps"] for lv in ok), default=None)
    best_ok = max((lv["docs_per_s"] for lv in ok), default=0.0)
    out["docs_per_s_with_p99_le_50ms"] = best_ok
    out["speedup_vs_plain_per_request_at_p99_le_50ms"] = best_ok / out["plain_per_request"]["docs_per_s"]
    log("seg_tokens %d: plain per-request %.0f docs/s (p50 %.1f ms); serve/plain %.1fx at p99 <= 50 ms (%.1fx at "
        "saturation); served ll identical to plain on the same requests: %s; served ll vs exact sampling %.2e"
        % (seg_tokens, out["plain_per_request"]["docs_per_s"], out["plain_per_request"]["p50_ms"],
           out["speedup_vs_plain_per_request_at_p99_le_50ms"], out["speedup_vs_plain_per_request"], same_ll,
           out["served_vs_exact_ll_rel_diff"]))
    return out


def serve_both():
    """Two operating points of the server on one model: the fidelity point
    (the autotuner's segmentation within 0.1% ll of exact sampling; lanes
    sweep and every baseline) and the latency point (seg_tokens = 128: every
    document over 128 tokens split, which shortens the per-wavefront chains
    that set the latency; its ll against exact sampling reported)."""
    setup = serve_setup()
    out = serve_bench(setup, ll_tol=1e-3)
    out["latency_point"] = serve_bench(setup, seg_tokens=128, lanes_sweep=(), cpu=False)
    return out


def big_k_bench(K=4096, n_docs=10000, sweeps=5, fold_docs=1000, fold_iters=20, cpu_budget_s=8.0):
    """K above 2048 (the device-memory-state samplers): the device-wide sweep
    on the large_k1000 generator cut to n_docs documents (tokens/s), and the
    fold-in of fold_docs of its documents (docs/s); the CPU oracle's
    partition sweep and fold-in on bounded samples of the same data."""
    from oracle import oracle as O
    from tomotopy_amd import _native, corpus, synth
    words, off, nw, _ = synth.workload("large_k1000", n_docs=n_docs + fold_docs)
    ws, offs, V, _ = corpus.prepare_arrays(words, off, nw)
    tw, toff = ws[: int(offs[n_docs])], offs[: n_docs + 1].astype(np.uint64)
    uw, uoff = ws[int(offs[n_docs]):], (offs[n_docs:] - offs[n_docs]).astype(np.uint64)
    n_valid = int(np.count_nonzero(tw < V))
    ctx = _native.Context(K, V, 0.05, 0.01, seed=77)
    ctx.upload_corpus(tw, toff)
    ctx.set_optim(0, 0)
    _, W = ctx.resolve_schedule(0, 0)
    ctx.train(2, workers=0)
    t = time.perf_counter()
    ctx.train(sweeps, workers=0)
    dt = (time.perf_counter() - t) / sweeps
    ms = ctx.kernel_stats()["ms"] / max(1, ctx.kernel_stats()["launches"])
    ctx.infer(uw[: int(uoff[50])], uoff[:51], 2)
    t = time.perf_counter()
    ctx.infer(uw, uoff, fold_iters)
    dtf = time.perf_counter() - t
    out = {"workload": "large_k1000 cut to %d docs" % n_docs, "K": K, "tokens": n_valid, "workers": W,
           "sweep_ms": dt * 1e3, "kernel_ms": ms, "tokens_per_s": n_valid / dt,
           "fold_in": {"docs": fold_docs, "iterations": fold_iters, "docs_per_s": fold_docs / dtf}}
    # CPU oracle: partition sweeps on a sample, fold-in of a document sample
    z, _, _, _ = ctx.download_state()
    thr = host_threads()
    sd = max(1, min(n_docs, 400))
    m = O.Model(K, V, 0.05, 0.01, 77, scan_order=1)
    m.set_optim(0, 0)
    m.set_corpus(tw[: int(toff[sd])], toff[: sd + 1])
    m.train(1, workers=thr, threads=thr)
    t, n = time.perf_counter(), 0
    while time.perf_counter() - t < cpu_budget_s / 2:
        m.train(1, workers=thr, threads=thr)
        n += 1
    stoks = int(np.count_nonzero(tw[: int(toff[sd])] < V))
    out["cpu_baseline"] = {"value": stoks * n / (time.perf_counter() - t), "unit": "tokens/s", "cores": thr,
                           "kind": "port", "sample": "first %d docs (%d tokens), partition with %d workers, %d sweeps"
                                                     % (sd, stoks, thr, n)}
    m2 = O.Model(K, V, 0.05, 0.01, 77, scan_order=1)
    m2.set_corpus(tw, toff, z)
    fd = 4 * thr
    from concurrent.futures import ThreadPoolExecutor

    def fold(d):  # the oracle releases the GIL
        b, e = int(uoff[d]), int(uoff[d + 1])
        m2.infer(uw[b:e], np.array([0, e - b], np.uint64), fold_iters, infer_seed=1)

    t = time.perf_counter()
    with ThreadPoolExecutor(thr) as ex:
        list(ex.map(fold, range(fd)))
    out["fold_in"]["cpu_baseline"] = {"value": fd / (time.perf_counter() - t), "unit": "docs/s", "cores": thr,
                                      "sample": "first %d documents, %d iterations, one per thread task"
                                                % (fd, fold_iters)}
    log("K=%d: device-wide %.1f ms/sweep (%d workers) = %.3g tokens/s, CPU oracle %.3g; fold-in %.0f docs/s, CPU %.0f"
        % (K, dt * 1e3, W, n_valid / dt, out["cpu_baseline"]["value"], fold_docs / dtf,
           out["fold_in"]["cpu_baseline"]["value"]))
    ctx.close()
    return out


def serve_pmi_bench(iterations=50, plain_requests=60, seed=7, rates=(50, 400)):
    """The server of a TermWeight.PMI model (run-column steps): the serve_k500
    workload with a K=500 PMI model; saturation throughput on 1 and 2 lanes,
    p50 / p99 at Poisson arrival rates, plain per-request lda_infer_w back to
    back, and served results against lda_infer_w on the same requests."""
    import threading
    from tomotopy_amd import _native, corpus, synth
    (tw, toff, nw), reqs_raw = synth.serve_workload(600)
    ws, offs, V, order = corpus.prepare_arrays(tw, toff, nw)
    reqs = [corpus.csr_from_arrays(w, o, order) for w, o in reqs_raw]
    vw = corpus.vocab_weights(ws, offs, V, "pmi")
    wts = corpus.term_weights(ws, offs, V, "pmi")
    ctx = _native.Context(500, V, 0.1, 0.01, seed=56, tw=2)
    ctx.upload_corpus(ws, offs, weights=wts)
    ctx.train(50, workers=0)
    rq = [(w, o, vw[w].astype(np.float32)) for w, o in reqs]
    ndocs = sum(len(o) - 1 for _, o, _ in rq)
    from tomotopy_amd import autotune
    allw = np.concatenate([w for w, _, _ in rq]).astype(np.uint32)
    allx = np.concatenate([x for _, _, x in rq]).astype(np.float32)
    alloff = np.zeros(ndocs + 1, np.uint64)
    alloff[1:] = np.cumsum(np.concatenate([np.diff(o.astype(np.int64)) for _, o, _ in rq]))
    tune = autotune.tune_infer(ctx, allw, alloff, iterations, infer_seed=seed, weights=allx)
    ctx.set_infer_options(tune["seg_tokens"], tune["seg_max"])
    log("serve PMI autotuner: seg_tokens %d seg_max %d" % (tune["seg_tokens"], tune["seg_max"]))
    out = {"workload": "serve_k500, TermWeight.PMI", "K": 500, "iterations": iterations, "saturation": {},
           "levels": [], "seg_tokens": tune["seg_tokens"], "seg_max": tune["seg_max"]}
    for lanes in (1, 2):
        with _native.Server(ctx, hbm_budget_bytes=8 << 30, iterations=iterations, infer_seed=seed,
                            lanes=lanes) as srv:
            srv.wait(srv.submit(*rq[0][:2], weights=rq[0][2]))
            t0 = time.perf_counter()
            ids = [srv.submit(w, o, weights=x) for w, o, x in rq]
            res = [srv.wait(i) for i in ids]
            dt = time.perf_counter() - t0
            lat = [r[2] for r in res]
            out["saturation"]["lanes_%d" % lanes] = {"docs_per_s": ndocs / dt, "p50_ms": _percentile(lat, 50),
                                                     "p99_ms": _percentile(lat, 99)}
            log("serve PMI saturation, %d lane(s): %.0f docs/s" % (lanes, ndocs / dt))
            if lanes == 1:
                same = True
                for j in (0, 17, 333):
                    th0, ll0 = ctx.infer(rq[j][0], rq[j][1], iterations, infer_seed=seed, weights=rq[j][2])[:2]
                    same = same and np.array_equal(res[j][0], th0) and np.array_equal(res[j][1], ll0)
                out["served_equals_infer_w"] = bool(same)
                rng = np.random.default_rng(seed)
                for qps in rates:
                    n = max(8, int(qps * 3.0))
                    gaps = rng.exponential(1.0 / qps, size=n)
                    pick = rng.integers(0, len(rq), size=n)
                    ids = [None] * n
                    t_start = time.perf_counter()

                    def client():
                        due = t_start
                        for j in range(n):
                            due += gaps[j]
                            d = due - time.perf_counter()
                            if d > 0:
                                time.sleep(d)
                            w, o, x = rq[pick[j]]
                            ids[j] = srv.submit(w, o, weights=x)

                    th = threading.Thread(target=client)
                    th.start()
                    lat, docs = [], 0
                    for j in range(n):
                        while ids[j] is None:
                            time.sleep(0.0005)
                        _, ll, ms_ = srv.wait(ids[j])
                        lat.append(ms_)
                        docs += len(ll)
                    th.join()
                    el = time.perf_counter() - t_start
                    out["levels"].append({"qps": qps, "requests": n, "docs_per_s": docs / el,
                                          "p50_ms": _percentile(lat, 50), "p99_ms": _percentile(lat, 99)})
                    log("serve PMI qps %d: %.0f docs/s, p50 %.1f ms, p99 %.1f ms"
                        % (qps, docs / el, _percentile(lat, 50), _percentile(lat, 99)))
    t0 = time.perf_counter()
    pd = 0
    for w, o, x in rq[:plain_requests]:
        ctx.infer(w, o, iterations, infer_seed=seed, weights=x)
        pd += len(o) - 1
    out["plain_per_request"] = {"requests": plain_requests, "docs_per_s": pd / (time.perf_counter() - t0)}
    log("serve PMI plain per-request lda_infer_w: %.0f docs/s" % out["plain_per_request"]["docs_per_s"])
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--docs", type=int, default=None, help="fewer documents (quick runs only)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--corpus-cache", default=None, help="npz file to reuse the generated corpus")
    ap.add_argument("--no-infer", action="store_true", help="skip the fold-in measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-pmi", action="store_true", help="skip the TermWeight.PMI training variant")
    ap.add_argument("--no-partition", action="store_true", help="skip the partition-schedule legs")
    ap.add_argument("--no-big-k", action="store_true", help="skip the K=4096 and PMI-server legs")
    ap.add_argument("--only", choices=["train", "infer", "infer_pmi", "serve", "probe", "big_k", "serve_pmi"],
                    default=None,
                    help="run one leg only (the JSON line then carries that leg's numbers)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))
    if args.only == "probe":
        print(json.dumps({"leg": "probe", "probe": bandwidth_probe()}), flush=True)
        return
    if args.only in ("infer", "infer_pmi", "serve", "big_k", "serve_pmi"):
        res = {"infer": infer_bench, "infer_pmi": infer_pmi_bench, "serve": serve_both, "big_k": big_k_bench,
               "serve_pmi": serve_pmi_bench}[args.only]()
        print(json.dumps({"metric": METRIC, "leg": args.only, args.only: res}), flush=True)
        return

    from tomotopy_amd import _native
    K, alpha, eta = 1000, 0.05, 0.01
    ws, offs, V, w = make_corpus(args.docs, args.corpus_cache)
    D = len(offs) - 1
    n_valid = int(np.count_nonzero(ws < V))
    ctx = _native.Context(K, V, alpha, eta, seed=1234)
    t = time.time()
    ctx.upload_corpus(ws, offs)
    log("upload + initial topics: %.2fs" % (time.time() - t))
    for i in range(args.warmup):
        ctx.train(1, workers=0)
        log("warmup step %d: %.2f ms kernel" % (i, ctx.kernel_stats()["ms"]))
    t0 = time.perf_counter()
    ctx.train(args.steps, workers=0)
    dt = time.perf_counter() - t0
    st = ctx.kernel_stats()
    flags = ctx.check_invariants()
    if flags.any():
        log("INVARIANT VIOLATION: %s" % flags.tolist())
        sys.exit(3)
    lld, llr = ctx.loglik()
    kern_ms = st["ms"] / max(1, st["launches"])
    narrow = bool(st["narrow"])
    kname = "sweep_device_kernel<16, false, %s>" % ("true" if narrow else "false")
    sb = sweep_bytes(ws, offs, V, K, narrow_hot(ws, V) if narrow else V)
    bytes_per_launch = sb["per_token"]
    achieved = bytes_per_launch / (kern_ms * 1e-3) / 1e9
    value = n_valid * args.steps / dt
    traffic, traffic_commit = pmc_traffic(kname)
    probe = bandwidth_probe()
    log("ll_per_word %.6f; kernel avg %.3f ms; wall %.3f ms/step; probe read %.0f GB/s, copy %.0f GB/s"
        % ((lld + llr) / n_valid, kern_ms, dt * 1e3 / args.steps, probe["read_GBps"], probe["copy_GBps"]))
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "tokens/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "large_k1000", "docs": D, "tokens": n_valid, "V": V, "K": K, "alpha": alpha,
                   "eta": eta, "scheme": "device-wide (workers=0)", "optim_interval": 10,
                   "narrow_columns": narrow, "int32_column_words": sb["hot_words"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_commit": traffic_commit,
                     "kernel": kname, "kernel_ms": kern_ms, "algorithmic_bytes_per_launch": bytes_per_launch,
                     "bytes_model": "per token 4K (int32 column) or 2K (u16 column) + 8, per document 8K",
                     "runs": sb["runs"], "algorithmic_bytes_per_launch_runs": sb["per_run"],
                     "frac_runs": sb["per_run"] / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "dense_int32_bytes_per_launch": sb["dense_int32"],
                     "dense_int32_GBps_equivalent": sb["dense_int32"] / (kern_ms * 1e-3) / 1e9,
                     "read_ceiling_GBps": probe["read_GBps"], "copy_ceiling_GBps": probe["copy_GBps"]},
        "ll_per_word": (lld + llr) / n_valid,
    }
    if traffic:
        out["roofline"]["traffic_GBps"] = traffic / (kern_ms * 1e-3) / 1e9
        out["roofline"]["traffic_frac_of_read_ceiling"] = out["roofline"]["traffic_GBps"] / probe["read_GBps"]
    if not args.no_pmi:
        ctx = None
        out["pmi_variant"] = pmi_bench(ws, offs, V, K, alpha, eta, args.steps, args.warmup)
    if not args.no_partition and args.only in (None, "train"):
        ctx = None
        out["partition_c3"] = partition_bench("large_k1000", ws, offs, V, K, alpha, sweeps=2)
        from tomotopy_amd import corpus as _corpus, synth as _synth
        rw, ro, rnw, _ = _synth.workload("readme_k100")
        rws, roffs, rV, _ = _corpus.prepare_arrays(rw, ro, rnw)
        out["partition_c2"] = partition_bench("readme_k100", rws, roffs, rV, 100, 0.1, sweeps=5,
                                              cpu_sweeps=0 if args.no_cpu_baseline else 5)
    if rank == 0 and not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(ws, offs, V, K, alpha, eta)
        out["cpu_baseline_w1"] = cpu_baseline(ws, offs, V, K, alpha, eta, budget_s=4.0, threads=1, max_docs=1000)
        log("CPU baseline: %.3g tokens/s with %d workers, %.3g with one (%s)"
            % (out["cpu_baseline"]["value"], out["cpu_baseline"]["cores"], out["cpu_baseline_w1"]["value"],
               out["cpu_baseline"]["cpu_model"]))
    if args.only == "train":
        pass
    elif not args.no_infer:
        ctx = None
        out["infer"] = infer_bench()
        out["infer_pmi"] = infer_pmi_bench()
    if not args.no_serve and args.only is None:
        out["serve"] = serve_both()
    if not args.no_big_k and args.only is None:
        out["big_k"] = big_k_bench()
        out["serve_pmi"] = serve_pmi_bench()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
