"""ctypes wrapper of the CPU oracle (oracle/lda_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product package never imports this module.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liblda_oracle.so")

TAGS = {"sweep": 0, "init": 1, "shuffle": 2, "infer": 3, "infinit": 4}

_c_p = ctypes.c_void_p
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    i32, i64, u32, u64, f32, f64 = (ctypes.c_int, ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint64,
                                    ctypes.c_float, ctypes.c_double)
    sigs = {
        "orc_philox4x32_10": (None, [_c_p, _c_p, _c_p]),
        "orc_token_uniform": (f32, [u64, u64, u32, u32]),
        "orc_tag": (u32, [i32]),
        "orc_device_kpl": (i32, [i32]),
        "orc_weights": (None, [i32, _c_p, _c_p, _c_p, f32, _c_p, f32, _c_p]),
        "orc_prefix_sum_tomoto": (None, [_c_p, i32]),
        "orc_prefix_sum_device": (None, [_c_p, i32]),
        "orc_search": (i32, [_c_p, i32, f32]),
        "orc_shuffled_order": (None, [u64, u32, _c_p]),
        "orc_markstein_mismatches": (u64, [u64, u64, f64, f64, f64, f64]),
        "orc_sample_frozen": (None, [i32, i32, i32, _c_p, f32, f32, _c_p, _c_p, _c_p, i64, _c_p, _c_p, _c_p, _c_p,
                                     _c_p, _c_p]),
        "orc_lgamma_lut": (f32, [f32]),
        "orc_digamma_lut": (f32, [f32]),
        "orc_digamma_approx": (f32, [f32]),
        "orc_create": (_c_p, [i32, i32, _c_p, f32, u64, i32]),
        "orc_destroy": (None, [_c_p]),
        "orc_set_optim": (None, [_c_p, u32, u32]),
        "orc_set_corpus": (None, [_c_p, _c_p, _c_p, i64, _c_p]),
        "orc_set_alpha": (None, [_c_p, _c_p]),
        "orc_get_alpha": (None, [_c_p, _c_p]),
        "orc_global_step": (u32, [_c_p]),
        "orc_train": (None, [_c_p, i32, i32, i32]),
        "orc_train_scheme": (None, [_c_p, i32, i32, i32, i32]),
        "orc_optimize_alpha": (None, [_c_p]),
        "orc_get_state": (None, [_c_p, _c_p, _c_p, _c_p, _c_p]),
        "orc_loglik": (None, [_c_p, i32, _c_p, _c_p]),
        "orc_infer": (None, [_c_p, _c_p, _c_p, i64, i32, u64, i32, _c_p, _c_p, _c_p]),
        "orc_infer_seg": (None, [_c_p, _c_p, _c_p, i64, i32, u64, i32, u32, i32, u32, _c_p, _c_p, _c_p]),
        "orc_seg_bounds": (i32, [_c_p, u64, u32, i32, _c_p]),
        "orc_infer_ll_bruteforce": (f64, [_c_p, _c_p, u64, _c_p]),
        "orc_set_corpus_w": (None, [_c_p, _c_p, _c_p, i64, _c_p, _c_p]),
        "orc_set_priors": (None, [_c_p, _c_p]),
        "orc_infer_together": (None, [_c_p, _c_p, _c_p, i64, i32, u64, i32, i32, i32, _c_p, _c_p, _c_p]),
        "orc_sample_from_discrete": (i32, [_c_p, i32, f32]),
        "orc_train_w": (None, [_c_p, i32, i32, i32, i32]),
        "orc_optimize_alpha_w": (None, [_c_p]),
        "orc_merge_w": (None, [_c_p]),
        "orc_get_state_w": (None, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p]),
        "orc_set_state_w": (None, [_c_p, _c_p, _c_p, _c_p, _c_p]),
        "orc_loglik_w": (None, [_c_p, _c_p, _c_p]),
        "orc_infer_gen": (None, [_c_p, _c_p, _c_p, i64, i32, u64, _c_p, i32, _c_p, _c_p, _c_p]),
        "orc_infer_gen_seg": (None, [_c_p, _c_p, _c_p, i64, i32, u64, _c_p, i32, u32, i32, _c_p, _c_p, _c_p]),
        "orc_infer_together_w": (None, [_c_p, _c_p, _c_p, i64, _c_p, i32, u64, i32, i32, i32, _c_p, _c_p, _c_p]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    load().orc_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def tag(name):
    return load().orc_tag(TAGS[name])


def token_uniform(seed, g, sweep, tag_name="sweep"):
    return load().orc_token_uniform(seed, g, sweep, tag(tag_name))


def device_kpl(K):
    return load().orc_device_kpl(K)


def weights(ndk, alpha, nkw, eta, nk, veta):
    K = len(ndk)
    out = np.zeros(K, dtype=np.float32)
    load().orc_weights(K, _p(np.ascontiguousarray(ndk, np.int32)), _p(np.ascontiguousarray(alpha, np.float32)),
                       _p(np.ascontiguousarray(nkw, np.int32)), float(eta),
                       _p(np.ascontiguousarray(nk, np.int32)), float(veta), _p(out))
    return out


def prefix_sum_tomoto(x):
    a = np.array(x, dtype=np.float32)
    load().orc_prefix_sum_tomoto(_p(a), len(a))
    return a


def prefix_sum_device(x, kpl):
    a = np.zeros(64 * kpl, dtype=np.float32)
    a[:len(x)] = x
    load().orc_prefix_sum_device(_p(a), kpl)
    return a


def search(cdf, K, u):
    return load().orc_search(_p(np.ascontiguousarray(cdf, np.float32)), K, float(u))


def sample_from_discrete(w, u):
    """tomoto's sampleFromDiscrete (subtractive scan) of one uniform."""
    w = np.ascontiguousarray(w, np.float32)
    return load().orc_sample_from_discrete(_p(w), len(w), float(u))


def shuffled_order(N, seed):
    out = np.zeros(N, dtype=np.uint32)
    load().orc_shuffled_order(N, seed, _p(out))
    return out


def markstein_mismatches(n, seed, amin, amax, bmin, bmax):
    return load().orc_markstein_mismatches(n, seed, amin, amax, bmin, bmax)


def sample_frozen(K, scan_order, alpha, eta, veta, ndk, nkw, nk, doc, word, z_old, u):
    kpl = device_kpl(K)
    n = len(doc)
    zn = np.zeros(n, dtype=np.uint16)
    tot = np.zeros(n, dtype=np.float32)
    load().orc_sample_frozen(K, scan_order, kpl, _p(np.ascontiguousarray(alpha, np.float32)), float(eta), float(veta),
                             _p(np.ascontiguousarray(ndk, np.int32)), _p(np.ascontiguousarray(nkw, np.int32)),
                             _p(np.ascontiguousarray(nk, np.int32)), n, _p(np.ascontiguousarray(doc, np.int64)),
                             _p(np.ascontiguousarray(word, np.uint32)), _p(np.ascontiguousarray(z_old, np.uint16)),
                             _p(np.ascontiguousarray(u, np.float32)), _p(zn), _p(tot))
    return zn, tot


def seg_bounds(words, seg_tokens, seg_max):
    """Segment starts (S + 1 offsets) of one document under the segmented fold-in rule."""
    w = np.ascontiguousarray(words, np.uint32)
    out = np.zeros(9, np.uint64)
    S = load().orc_seg_bounds(_p(w), len(w), int(seg_tokens), int(seg_max), _p(out))
    return out[:S + 1]


def lgamma_lut(x):
    return load().orc_lgamma_lut(float(x))


def digamma_lut(x):
    return load().orc_digamma_lut(float(x))


def digamma_approx(x):
    return load().orc_digamma_approx(float(x))


class Model:
    """The oracle's plain-LDA model (tomoto semantics, Philox stream)."""

    def __init__(self, K, V, alpha, eta, seed, scan_order=1):
        self.K, self.V = K, V
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(alpha, np.float32), (K,)), dtype=np.float32)
        self._h = load().orc_create(K, V, _p(a), float(eta), int(seed) & (2**64 - 1), int(scan_order))
        self.D = 0
        self.N = 0

    def __del__(self):
        if getattr(self, "_h", None):
            load().orc_destroy(self._h)
            self._h = None

    def set_optim(self, interval, burn_in=0):
        load().orc_set_optim(self._h, interval, burn_in)

    def set_corpus(self, words, off, z=None):
        self._words = np.ascontiguousarray(words, np.uint32)
        self._off = np.ascontiguousarray(off, np.uint64)
        zz = None if z is None else np.ascontiguousarray(z, np.uint16)
        self.D = len(off) - 1
        self.N = int(off[-1])
        load().orc_set_corpus(self._h, _p(self._words), _p(self._off), self.D, _p(zz))

    def set_priors(self, eta_kv):
        """Word priors: eta_kv (V x K); call before set_corpus."""
        self._eta_kv = np.ascontiguousarray(eta_kv, np.float32).reshape(self.V, self.K)
        load().orc_set_priors(self._h, _p(self._eta_kv))

    def set_alpha(self, alpha):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(alpha, np.float32), (self.K,)), dtype=np.float32)
        load().orc_set_alpha(self._h, _p(a))

    def alpha(self):
        a = np.zeros(self.K, np.float32)
        load().orc_get_alpha(self._h, _p(a))
        return a

    def train(self, iterations, workers=1, threads=1, scheme=3):
        """scheme: 1 none, 2 copy_merge, 3 partition (tomotopy.ParallelScheme values)."""
        load().orc_train_scheme(self._h, iterations, workers, threads, scheme)

    def optimize_alpha(self):
        load().orc_optimize_alpha(self._h)

    @property
    def global_step(self):
        return load().orc_global_step(self._h)

    def state(self):
        z = np.zeros(self.N, np.uint16)
        ndk = np.zeros((self.D, self.K), np.int32)
        nkw = np.zeros((self.V, self.K), np.int32)
        nk = np.zeros(self.K, np.int32)
        load().orc_get_state(self._h, _p(z), _p(ndk), _p(nkw), _p(nk))
        return z, ndk, nkw, nk

    def loglik(self, mode=0):
        a = ctypes.c_double()
        b = ctypes.c_double()
        load().orc_loglik(self._h, mode, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def infer(self, words, off, iters, infer_seed=0, mode=0, seg_tokens=0, seg_max=1, seg_sync=64):
        """Fold-in; seg_tokens / seg_max select the segmented rule (0: sequential);
        seg_sync: the segments merge their n_dk changes after every seg_sync tokens
        of each segment (64: the integer device kernel's rule), or once per
        iteration (0)."""
        words = np.ascontiguousarray(words, np.uint32)
        off = np.ascontiguousarray(off, np.uint64)
        D = len(off) - 1
        theta = np.zeros((D, self.K), np.float32)
        ll = np.zeros(D, np.float64)
        z = np.zeros(int(off[-1]), np.uint16)
        load().orc_infer_seg(self._h, _p(words), _p(off), D, iters, int(infer_seed) & (2**64 - 1), mode,
                             int(seg_tokens), int(seg_max), int(seg_sync), _p(theta), _p(ll), _p(z))
        return theta, ll, z

    def infer_gen(self, words, off, iters, infer_seed=0, weights=None, faithful=False, seg_tokens=0, seg_max=1):
        """Fold-in of a term-weighted model (weights: per-token float32) or of
        a model with word priors; returns (theta, ll, z).  faithful keeps the
        full K x V private copy per document (faithful=2: and evaluates the LL
        as tomoto does, with its lgamma table, for timing tomoto's form);
        seg_tokens / seg_max select the segmented rule (0: sequential)."""
        words = np.ascontiguousarray(words, np.uint32)
        off = np.ascontiguousarray(off, np.uint64)
        D = len(off) - 1
        theta = np.zeros((D, self.K), np.float32)
        ll = np.zeros(D, np.float64)
        z = np.zeros(int(off[-1]), np.uint16)
        w = None if weights is None else np.ascontiguousarray(weights, np.float32)
        load().orc_infer_gen_seg(self._h, _p(words), _p(off), D, iters, int(infer_seed) & (2**64 - 1), _p(w),
                                 int(faithful), int(seg_tokens), int(seg_max), _p(theta), _p(ll), _p(z))
        return theta, ll, z

    def infer_together(self, words, off, iters, infer_seed=0, workers=1, scheme=1, threads=1, weights=None):
        """Fold-in with together=True: returns (theta, ll, z).  weights: the
        tokens' weights of a term-weighted model."""
        words = np.ascontiguousarray(words, np.uint32)
        off = np.ascontiguousarray(off, np.uint64)
        D = len(off) - 1
        theta = np.zeros((D, self.K), np.float32)
        z = np.zeros(int(off[-1]), np.uint16)
        ll = ctypes.c_double()
        seed = int(infer_seed) & (2**64 - 1)
        if weights is None:
            load().orc_infer_together(self._h, _p(words), _p(off), D, iters, seed, workers,
                                      scheme, threads, _p(theta), ctypes.byref(ll), _p(z))
        else:
            w = np.ascontiguousarray(weights, np.float32)
            load().orc_infer_together_w(self._h, _p(words), _p(off), D, _p(w), iters, seed, workers,
                                        scheme, threads, _p(theta), ctypes.byref(ll), _p(z))
        return theta, ll.value, z

    def infer_ll_bruteforce(self, words, z):
        words = np.ascontiguousarray(words, np.uint32)
        z = np.ascontiguousarray(z, np.uint16)
        return load().orc_infer_ll_bruteforce(self._h, _p(words), len(words), _p(z))


class WeightedModel(Model):
    """The oracle's term-weighted model (TermWeight IDF / PMI): float counts,
    clamped removal, weighted merge (LDAModel.hpp:287-299, 565-579)."""

    def set_corpus(self, words, off, weights, z=None):
        self._words = np.ascontiguousarray(words, np.uint32)
        self._off = np.ascontiguousarray(off, np.uint64)
        self._wt = np.ascontiguousarray(weights, np.float32)
        zz = None if z is None else np.ascontiguousarray(z, np.uint16)
        self.D = len(off) - 1
        self.N = int(off[-1])
        load().orc_set_corpus_w(self._h, _p(self._words), _p(self._off), self.D, _p(zz), _p(self._wt))

    def train(self, iterations, workers=1, threads=1, scheme=3):
        """scheme: 1 none, 2 copy_merge, 3 partition."""
        load().orc_train_w(self._h, iterations, workers, threads, scheme)

    def optimize_alpha(self):
        load().orc_optimize_alpha_w(self._h)

    def merge(self):
        load().orc_merge_w(self._h)

    def state(self):
        z = np.zeros(self.N, np.uint16)
        ndk = np.zeros((self.D, self.K), np.float32)
        nkw = np.zeros((self.V, self.K), np.float32)
        nk = np.zeros(self.K, np.float32)
        docw = np.zeros(self.D, np.float32)
        load().orc_get_state_w(self._h, _p(z), _p(ndk), _p(nkw), _p(nk), _p(docw))
        return z, ndk, nkw, nk, docw

    def set_state(self, z=None, ndk=None, nkw=None, nk=None):
        c = lambda a, t: None if a is None else np.ascontiguousarray(a, t)
        z, ndk, nkw, nk = c(z, np.uint16), c(ndk, np.float32), c(nkw, np.float32), c(nk, np.float32)
        load().orc_set_state_w(self._h, _p(z), _p(ndk), _p(nkw), _p(nk))

    def loglik(self, mode=0):
        if mode != 0:
            raise ValueError("the term-weighted log-likelihood is evaluated in FP64 only")
        a = ctypes.c_double()
        b = ctypes.c_double()
        load().orc_loglik_w(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value


def term_weights(words, off, V, scheme):
    """Token weights of TermWeight IDF ("idf") or PMI ("pmi"), restated in
    plain loops from LDAModel.hpp:1096-1121 (vocabulary weights) and 800-809
    (per-token weights).  words: per-document token ids (ids >= V are removed
    words: weight 0, but they count in the document length and in the total
    collection frequency).  IDF: w = float(log(D / df[v])).  PMI: vw[v] =
    float(cf[v] / totCf), w = max(float(log(float(tf_d[v]) / vw[v] / float(len_d))), 0)
    with the quotient in float and the logarithm taken in double."""
    D = len(off) - 1
    out = np.zeros(len(words), np.float32)
    if scheme == "idf":
        df = [0] * V
        for d in range(D):
            for v in set(int(x) for x in words[int(off[d]):int(off[d + 1])]):
                if v < V:
                    df[v] += 1
        vw = [np.float32(math.log(D / df[v])) if df[v] else np.float32(np.inf) for v in range(V)]
        for g, v in enumerate(words):
            if int(v) < V:
                out[g] = vw[int(v)]
        return out
    if scheme != "pmi":
        raise ValueError(scheme)
    cf = [0] * V
    tot = 0
    for v in words:
        tot += 1
        if int(v) < V:
            cf[int(v)] += 1
    vw = [np.float32(cf[v] / tot) for v in range(V)]
    for d in range(D):
        b, e = int(off[d]), int(off[d + 1])
        tf = {}
        for v in words[b:e]:
            if int(v) < V:
                tf[int(v)] = tf.get(int(v), 0) + 1
        n = np.float32(e - b)
        for g in range(b, e):
            v = int(words[g])
            if v < V:
                q = np.float32(np.float32(tf[v]) / vw[v]) / n
                out[g] = max(np.float32(math.log(float(q))), np.float32(0.0))
    return out
