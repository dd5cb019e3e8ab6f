"""ctypes binding of the C ABI in include/tomotopy_amd.h.

The HIP library is the only compute path.  If it is missing, or no HIP device
is visible, every entry point raises; nothing falls back to a CPU sampler.
"""
import ctypes
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libtomotopy_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "tomotopy_amd.h")

LDA_OK = 0
LDA_ERR_INVALID_ARGUMENT = -1
LDA_ERR_STATE = -2
LDA_ERR_DEVICE = -3
LDA_ERR_OUT_OF_MEMORY = -4
LDA_ERR_UNSUPPORTED = -5
LDA_ERR_TRAINING = -6

LDA_MAX_K = 32767

_c_p = ctypes.c_void_p
_i32 = ctypes.c_int
_i64 = ctypes.c_int64
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_f32 = ctypes.c_float
_dp = ctypes.POINTER(ctypes.c_double)

# name -> argtypes (restype is int for all but lda_last_error)
_SIGS = {
    "lda_last_error": [],
    "lda_version": [],
    "lda_device_count": [_c_p],
    "lda_ctx_create": [_i32, _i32, _c_p, _f32, _i32, _u64, _i32, _c_p],
    "lda_ctx_destroy": [_c_p],
    "lda_ctx_copy": [_c_p, _c_p],
    "lda_debug_fail_at_step": [_c_p, _i64],
    "lda_upload_corpus": [_c_p, _c_p, _c_p, _i64, _c_p, _c_p],
    "lda_set_counts": [_c_p, _c_p, _c_p],
    "lda_set_counts_f": [_c_p, _c_p, _c_p],
    "lda_set_word_priors": [_c_p, ctypes.c_int64, _c_p, _c_p],
    "lda_selftest_reciprocal": [_i32, _i32, _c_p],
    "lda_probe_bandwidth": [_i32, _u64, _i32, _c_p],
    "lda_topic_word_dist": [_c_p, _i32, _i32, _c_p],
    "lda_count_by_topic": [_c_p, _c_p],
    "lda_doc_topic_dist": [_c_p, _i64, _i32, _c_p],
    "lda_infer_together": [_c_p, _c_p, _c_p, ctypes.c_int64, _i32, _u64, _i32, _i32, _c_p, _c_p, _c_p],
    "lda_infer_together_w": [_c_p, _c_p, _c_p, _i64, _c_p, _i32, _u64, _i32, _i32, _c_p, _c_p, _c_p],
    "lda_set_alpha": [_c_p, _c_p],
    "lda_get_alpha": [_c_p, _c_p],
    "lda_set_optim": [_c_p, _u32, _u32],
    "lda_get_global_step": [_c_p, _c_p],
    "lda_set_global_step": [_c_p, _u64],
    "lda_train": [_c_p, _i32, _i32, _i32],
    "lda_sweep": [_c_p, _i32, _i32],
    "lda_resolve_schedule": [_c_p, _i32, _i32, _c_p, _c_p],
    "lda_set_device_schedule": [_c_p, _i32, _i32],
    "lda_optimize_alpha": [_c_p],
    "lda_loglik": [_c_p, _c_p, _c_p],
    "lda_check_invariants": [_c_p, _c_p],
    "lda_download_state": [_c_p, _c_p, _c_p, _c_p, _c_p],
    "lda_get_dims": [_c_p, _c_p],
    "lda_download_state_f": [_c_p, _c_p, _c_p, _c_p, _c_p],
    "lda_get_weighted_n": [_c_p, _c_p],
    "lda_sample_frozen": [_c_p, _i64, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p],
    "lda_infer": [_c_p, _c_p, _c_p, _i64, _i32, _u64, _i32, _c_p, _c_p, _c_p],
    "lda_infer_noalloc": [_c_p, _c_p, _c_p, _i64, _i32, _u64, _c_p, _c_p, _c_p],
    "lda_infer_w": [_c_p, _c_p, _c_p, _i64, _c_p, _i32, _u64, _i32, _c_p, _c_p, _c_p],
    "lda_kernel_stats": [_c_p, _c_p],
    "lda_set_narrow_columns": [_c_p, _i32],
    "lda_set_atomic_group": [_c_p, _i32],
    "lda_set_infer_options": [_c_p, _u32, _i32],
    "lda_get_infer_options": [_c_p, _c_p, _c_p, _c_p],
    "lda_serve_start": [_c_p, _u64, _i32, _i32, _i64, _i32, _u64, _c_p],
    "lda_serve_submit": [_c_p, _c_p, _c_p, _i64, _c_p],
    "lda_serve_submit_w": [_c_p, _c_p, _c_p, _i64, _c_p, _c_p],
    "lda_serve_wait": [_c_p, _u64, _i32, _c_p, _c_p, _c_p],
    "lda_serve_stats": [_c_p, _c_p],
    "lda_serve_stop": [_c_p],
}

_lib = None


class NativeLibraryMissing(RuntimeError):
    pass


def header_symbols():
    """Function names declared in include/tomotopy_amd.h."""
    with open(HEADER_PATH) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(lda_\w+)\s*\(", text, re.M)))


def load():
    """Load the HIP library (raises NativeLibraryMissing when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryMissing(
            "tomotopy_amd HIP library not found at %s; build it with `python -m tomotopy_amd.build`" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_char_p if name == "lda_last_error" else ctypes.c_int
    _lib = lib
    return lib


def _check(rc):
    if rc == LDA_OK:
        return
    msg = load().lda_last_error().decode("utf-8", "replace")
    if rc == LDA_ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if rc == LDA_ERR_OUT_OF_MEMORY:
        raise MemoryError(msg)
    if rc == LDA_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == LDA_ERR_TRAINING:
        raise RuntimeError("Train failed: " + msg)
    raise RuntimeError(msg)


def device_count():
    n = ctypes.c_int(0)
    _check(load().lda_device_count(ctypes.byref(n)))
    return n.value


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Context:
    """One model's device state (n_kw, n_k, n_dk, z, corpus) on one GPU."""

    def __init__(self, k, v, alpha, eta, seed, tw=0, device=0):
        lib = load()
        alpha = np.ascontiguousarray(np.broadcast_to(np.asarray(alpha, dtype=np.float32), (k,)), dtype=np.float32)
        h = ctypes.c_void_p()
        _check(lib.lda_ctx_create(int(k), int(v), _ptr(alpha), float(eta), int(tw), int(seed) & (2**64 - 1),
                                  int(device), ctypes.byref(h)))
        self._h = h
        self.k = int(k)
        self.v = int(v)
        self.tw = int(tw)
        self.n_docs = 0
        self.n_tokens = 0

    def copy(self):
        """A deep copy of this context's device state (lda_ctx_copy)."""
        h = ctypes.c_void_p()
        _check(load().lda_ctx_copy(self._h, ctypes.byref(h)))
        new = Context.__new__(Context)
        new._h = h
        new.k, new.v, new.tw = self.k, self.v, self.tw
        new.n_docs, new.n_tokens = self.n_docs, self.n_tokens
        return new

    def topic_word_dist(self, topic, normalize=True):
        """(n_kw + eta) / (n_k + V*eta) of one topic over the words, read out on the device."""
        out = np.zeros(self.v, dtype=np.float32)
        _check(load().lda_topic_word_dist(self._h, int(topic), int(bool(normalize)), _ptr(out)))
        return out

    def count_by_topic(self):
        """Tokens assigned to each topic (valid tokens only), counted on the device."""
        out = np.zeros(self.k, dtype=np.uint64)
        _check(load().lda_count_by_topic(self._h, _ptr(out)))
        return out

    def doc_topic_dist(self, doc, normalize=True):
        """(n_dk + alpha) / (weight sum + sum alpha) of training document `doc`."""
        out = np.zeros(self.k, dtype=np.float32)
        _check(load().lda_doc_topic_dist(self._h, int(doc), int(bool(normalize)), _ptr(out)))
        return out

    def fail_at_step(self, step):
        """Test hook: the training iteration starting at global step `step` fails (-1 clears)."""
        _check(load().lda_debug_fail_at_step(self._h, int(step)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            load().lda_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_corpus(self, words, doc_off, z=None, weights=None):
        words = np.ascontiguousarray(words, dtype=np.uint32)
        doc_off = np.ascontiguousarray(doc_off, dtype=np.uint64)
        if z is not None:
            z = np.ascontiguousarray(z, dtype=np.uint16)
        if weights is not None:
            weights = np.ascontiguousarray(weights, dtype=np.float32)
        D = len(doc_off) - 1
        _check(load().lda_upload_corpus(self._h, _ptr(words), _ptr(doc_off), D, _ptr(z), _ptr(weights)))
        self.n_docs = D
        self.n_tokens = int(doc_off[-1]) if D >= 0 and len(doc_off) else 0

    def set_counts(self, nkw, nk):
        nkw = np.ascontiguousarray(nkw, dtype=np.int32)
        nk = np.ascontiguousarray(nk, dtype=np.int32)
        _check(load().lda_set_counts(self._h, _ptr(nkw), _ptr(nk)))

    def set_word_priors(self, word_ids, priors):
        """priors: len(word_ids) x K; call before upload_corpus."""
        word_ids = np.ascontiguousarray(word_ids, dtype=np.uint32)
        priors = np.ascontiguousarray(priors, dtype=np.float32).reshape(len(word_ids), self.k)
        _check(load().lda_set_word_priors(self._h, len(word_ids), _ptr(word_ids), _ptr(priors)))

    def set_counts_f(self, nkw, nk):
        nkw = np.ascontiguousarray(nkw, dtype=np.float32)
        nk = np.ascontiguousarray(nk, dtype=np.float32)
        _check(load().lda_set_counts_f(self._h, _ptr(nkw), _ptr(nk)))

    def set_alpha(self, alpha):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(alpha, dtype=np.float32), (self.k,)), dtype=np.float32)
        _check(load().lda_set_alpha(self._h, _ptr(a)))

    def get_alpha(self):
        a = np.zeros(self.k, dtype=np.float32)
        _check(load().lda_get_alpha(self._h, _ptr(a)))
        return a

    def set_optim(self, optim_interval, burn_in):
        _check(load().lda_set_optim(self._h, int(optim_interval), int(burn_in)))

    @property
    def global_step(self):
        s = ctypes.c_uint64(0)
        _check(load().lda_get_global_step(self._h, ctypes.byref(s)))
        return s.value

    @global_step.setter
    def global_step(self, v):
        _check(load().lda_set_global_step(self._h, int(v)))

    def train(self, iterations, workers=0, scheme=0):
        _check(load().lda_train(self._h, int(iterations), int(workers), int(scheme)))

    def resolve_schedule(self, workers=0, scheme=0):
        """(scheme, workers) that train(workers, scheme) runs: scheme 1 none, 2 copy_merge,
        3 partition, 0 the device-wide schedule (workers = its workgroups)."""
        so, wo = ctypes.c_int(0), ctypes.c_int(0)
        _check(load().lda_resolve_schedule(self._h, int(workers), int(scheme), ctypes.byref(so), ctypes.byref(wo)))
        return so.value, wo.value

    def set_device_schedule(self, max_workers=-1, waves_per_worker=8):
        """Concurrency of the device-wide schedule: at most max_workers workgroups (0: all
        resident), each of waves_per_worker wavefronts sharing one private n_k; max_workers < 0:
        the default (at most ceil(D / 8) documents in flight)."""
        _check(load().lda_set_device_schedule(self._h, int(max_workers), int(waves_per_worker)))

    def sweep(self, workers=0, scheme=0):
        _check(load().lda_sweep(self._h, int(workers), int(scheme)))

    def optimize_alpha(self):
        _check(load().lda_optimize_alpha(self._h))

    def loglik(self):
        a = ctypes.c_double(0.0)
        b = ctypes.c_double(0.0)
        _check(load().lda_loglik(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def check_invariants(self):
        f = np.zeros(8, dtype=np.int64)
        _check(load().lda_check_invariants(self._h, _ptr(f)))
        return f

    def dims(self):
        d = np.zeros(6, dtype=np.int64)
        _check(load().lda_get_dims(self._h, _ptr(d)))
        return d

    def download_state(self, z=True, ndk=True, nkw=True, nk=True):
        K, V, D, N = (int(x) for x in self.dims()[:4])
        zz = np.zeros(N, dtype=np.uint16) if z else None
        a = np.zeros((D, K), dtype=np.int32) if ndk else None
        b = np.zeros((V, K), dtype=np.int32) if nkw else None
        c = np.zeros(K, dtype=np.int32) if nk else None
        _check(load().lda_download_state(self._h, _ptr(zz), _ptr(a), _ptr(b), _ptr(c)))
        return zz, a, b, c

    def download_state_f(self):
        """Counts of a term-weighted model (float32)."""
        K, V, D, N = (int(x) for x in self.dims()[:4])
        zz = np.zeros(N, dtype=np.uint16)
        a = np.zeros((D, K), dtype=np.float32)
        b = np.zeros((V, K), dtype=np.float32)
        c = np.zeros(K, dtype=np.float32)
        _check(load().lda_download_state_f(self._h, _ptr(zz), _ptr(a), _ptr(b), _ptr(c)))
        return zz, a, b, c

    def weighted_n(self):
        x = ctypes.c_double(0.0)
        _check(load().lda_get_weighted_n(self._h, ctypes.byref(x)))
        return x.value

    def sample_frozen(self, doc, word, z_old, u):
        doc = np.ascontiguousarray(doc, dtype=np.int64)
        word = np.ascontiguousarray(word, dtype=np.uint32)
        z_old = np.ascontiguousarray(z_old, dtype=np.uint16)
        u = np.ascontiguousarray(u, dtype=np.float32)
        n = len(doc)
        zn = np.zeros(n, dtype=np.uint16)
        tot = np.zeros(n, dtype=np.float32)
        _check(load().lda_sample_frozen(self._h, n, _ptr(doc), _ptr(word), _ptr(z_old), _ptr(u), _ptr(zn), _ptr(tot)))
        return zn, tot

    def infer(self, words, doc_off, iterations, infer_seed=0, mode=0, want_z=False, weights=None):
        """Fold-in (lda_infer; lda_infer_w with the token weights of a
        term-weighted model)."""
        words = np.ascontiguousarray(words, dtype=np.uint32)
        doc_off = np.ascontiguousarray(doc_off, dtype=np.uint64)
        D = len(doc_off) - 1
        theta = np.zeros((max(D, 0), self.k), dtype=np.float32)
        ll = np.zeros(max(D, 0), dtype=np.float64)
        z = np.zeros(int(doc_off[-1]) if D > 0 else 0, dtype=np.uint16) if want_z else None
        seed = int(infer_seed) & (2**64 - 1)
        if weights is None:
            _check(load().lda_infer(self._h, _ptr(words), _ptr(doc_off), D, int(iterations), seed,
                                    int(mode), _ptr(theta), _ptr(ll), _ptr(z)))
        else:
            wts = np.ascontiguousarray(weights, dtype=np.float32)
            if len(wts) != len(words):
                raise ValueError("weights must have one entry per token")
            _check(load().lda_infer_w(self._h, _ptr(words), _ptr(doc_off), D, _ptr(wts), int(iterations), seed,
                                      int(mode), _ptr(theta), _ptr(ll), _ptr(z)))
        return theta, ll, z

    def infer_noalloc(self, words, doc_off, iterations, infer_seed=0, want_z=False):
        """Batched fold-in staged in the library's fixed device scratch (lda_infer_noalloc):
        no device allocation; same results as infer(mode=0)."""
        words = np.ascontiguousarray(words, dtype=np.uint32)
        doc_off = np.ascontiguousarray(doc_off, dtype=np.uint64)
        D = len(doc_off) - 1
        theta = np.zeros((max(D, 0), self.k), dtype=np.float32)
        ll = np.zeros(max(D, 0), dtype=np.float64)
        z = np.zeros(int(doc_off[-1]) if D > 0 else 0, dtype=np.uint16) if want_z else None
        _check(load().lda_infer_noalloc(self._h, _ptr(words), _ptr(doc_off), D, int(iterations),
                                        int(infer_seed) & (2**64 - 1), _ptr(theta), _ptr(ll), _ptr(z)))
        return theta, ll, z

    def infer_together(self, words, doc_off, iterations, infer_seed=0, workers=0, scheme=0, weights=None):
        """Fold-in with together=True: (theta D x K, ll for all documents, z).
        weights: the tokens' weights of a term-weighted model."""
        words = np.ascontiguousarray(words, dtype=np.uint32)
        doc_off = np.ascontiguousarray(doc_off, dtype=np.uint64)
        D = len(doc_off) - 1
        theta = np.zeros((max(D, 0), self.k), dtype=np.float32)
        z = np.zeros(int(doc_off[-1]) if len(doc_off) else 0, dtype=np.uint16)
        ll = ctypes.c_double(0.0)
        seed = int(infer_seed) & (2**64 - 1)
        if weights is None:
            _check(load().lda_infer_together(self._h, _ptr(words), _ptr(doc_off), D, int(iterations), seed,
                                             int(workers), int(scheme), _ptr(theta), ctypes.byref(ll), _ptr(z)))
        else:
            wts = np.ascontiguousarray(weights, dtype=np.float32)
            if len(wts) != len(words):
                raise ValueError("weights must have one entry per token")
            _check(load().lda_infer_together_w(self._h, _ptr(words), _ptr(doc_off), D, _ptr(wts), int(iterations),
                                               seed, int(workers), int(scheme), _ptr(theta), ctypes.byref(ll),
                                               _ptr(z)))
        return theta, ll.value, z

    def set_infer_options(self, seg_tokens=0, seg_max=1):
        """Segmented fold-in: documents longer than seg_tokens tokens are sampled by
        up to seg_max wavefronts (0: every document sequentially, as tomotopy)."""
        _check(load().lda_set_infer_options(self._h, int(seg_tokens), int(seg_max)))

    def infer_options(self):
        t = ctypes.c_uint32(0)
        m = ctypes.c_int(0)
        sup = ctypes.c_int(0)
        _check(load().lda_get_infer_options(self._h, ctypes.byref(t), ctypes.byref(m), ctypes.byref(sup)))
        return {"seg_tokens": t.value, "seg_max": m.value, "seg_supported": sup.value}

    def kernel_stats(self):
        s = np.zeros(4, dtype=np.float64)
        _check(load().lda_kernel_stats(self._h, _ptr(s)))
        return {"ms": float(s[0]), "launches": int(s[1]), "tokens": float(s[2]), "narrow": bool(s[3])}

    def set_narrow_columns(self, enable=True):
        """u16 n_kw columns of the less frequent words in the device-wide sweep (on by default)."""
        _check(load().lda_set_narrow_columns(self._h, int(bool(enable))))

    def set_atomic_group(self, tokens=0):
        """Tokens per group of shared n_kw atomics in the device-wide sweeps (0: chosen per launch)."""
        _check(load().lda_set_atomic_group(self._h, int(tokens)))


class Server:
    """Continuous-batching fold-in server on a Context's frozen model.

    submit() copies a request and returns its id at once; wait() blocks (the
    GIL is released inside the C call) and returns (theta, ll, latency_ms).
    Results equal Context.infer(mode=0) with the same infer_seed bit for bit
    (for a term-weighted model, submit the token weights as for infer).
    The context must not be trained or changed while the server runs.
    """

    def __init__(self, ctx, hbm_budget_bytes=1 << 30, iterations=50, step_iters=50, step_tokens=1 << 30,
                 infer_seed=0, lanes=2):
        self._ctx = ctx  # keeps the model alive while the server runs
        self.k = ctx.k
        self.iterations = int(iterations)
        h = ctypes.c_void_p()
        _check(load().lda_serve_start(ctx._h, int(hbm_budget_bytes), int(iterations), int(step_iters),
                                      int(step_tokens), int(lanes), int(infer_seed) & (2**64 - 1), ctypes.byref(h)))
        self._h = h
        self._ndocs = {}

    def submit(self, words, doc_off, weights=None):
        """weights: the tokens' weights, for the server of a term-weighted model."""
        words = np.ascontiguousarray(words, dtype=np.uint32)
        doc_off = np.ascontiguousarray(doc_off, dtype=np.uint64)
        D = len(doc_off) - 1
        rid = ctypes.c_uint64(0)
        if weights is None:
            _check(load().lda_serve_submit(self._h, _ptr(words), _ptr(doc_off), D, ctypes.byref(rid)))
        else:
            wts = np.ascontiguousarray(weights, dtype=np.float32)
            if len(wts) != len(words):
                raise ValueError("weights must have one entry per token")
            _check(load().lda_serve_submit_w(self._h, _ptr(words), _ptr(doc_off), D, _ptr(wts), ctypes.byref(rid)))
        self._ndocs[rid.value] = D
        return rid.value

    def wait(self, request_id, timeout_ms=-1):
        if request_id not in self._ndocs:
            raise ValueError("unknown request %r" % (request_id,))
        D = self._ndocs[request_id]
        theta = np.zeros((D, self.k), dtype=np.float32)
        ll = np.zeros(D, dtype=np.float64)
        lat = ctypes.c_double(0.0)
        _check(load().lda_serve_wait(self._h, int(request_id), int(timeout_ms), _ptr(theta), _ptr(ll),
                                     ctypes.byref(lat)))
        del self._ndocs[request_id]
        return theta, ll, lat.value

    def stats(self):
        s = np.zeros(12, dtype=np.float64)
        _check(load().lda_serve_stats(self._h, _ptr(s)))
        keys = ("steps", "busy_ms", "docs_done", "requests_done", "token_iters", "active_docs", "queued_requests",
                "arena_tokens", "gpu_ms", "host_ms", "max_step_ms", "docs_per_step")
        return dict(zip(keys, (float(x) for x in s)))

    def stop(self):
        if getattr(self, "_h", None) and self._h.value:
            load().lda_serve_stop(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass


def selftest_reciprocal(exp_shift, device=0):
    """Mismatches of the device reciprocal against IEEE division over one binade."""
    n = ctypes.c_uint64(0)
    _check(load().lda_selftest_reciprocal(int(exp_shift), int(device), ctypes.byref(n)))
    return int(n.value)


def probe_bandwidth(bytes_=1 << 30, reps=20, device=0):
    """(streaming read GB/s, copy GB/s) of the device (lda_probe_bandwidth)."""
    out = np.zeros(2, dtype=np.float64)
    _check(load().lda_probe_bandwidth(int(device), int(bytes_), int(reps), _ptr(out)))
    return float(out[0]), float(out[1])
