"""tomoto's binary model file layout (LDAModel.save / saves / load / loads).

Restated from the reference's serializer (src/Utils/serializer.hpp,
serializer.cpp:120-164) and the members it writes for an LDA model:

  key   b"LDA\\0\\0", then b"one\\0\\0" | b"idf\\0\\0" | b"pmi\\0\\0"
        (to_keyz of tmid / twid, TopicModel.hpp:335-359, LDAModel.hpp:120-150)
  tagged list: dict, vocabCf, vocabDf, realV, globalStep, extra
  tagged list: vocabWeights, alpha, alphas, eta, K, etaByWord, burnIn,
        optimInterval (LDAModel.hpp:957-960)
  n_k   u32 rows = K, u32 cols = 1, K counts (ModelStateLDA, LDAModel.hpp:58-69)
  n_kw  u32 rows = K, u32 cols = realV, K*realV counts column-major, i.e.
        the K counts of word 0, then of word 1, ... (ShareableMatrix, LDA.h:80-95)
  docs  u32 count, then per document the DocumentBase tagged list (weight,
        words, wOrder, rawStr, origWordPos, origWordLen, docUid,
        TopicModel.hpp:127-131) and the DocumentLDA tagged list (Zs,
        wordWeights, LDAModel.cpp:5-6); an empty list when full=False.

A tagged list is a run of entries, each a 24-byte header
  u32 "TPTK" (0x4b545054), u32 version 0x00010001, u64 totsize,
  u32 keysize, u32 trailing_cnt (entries still to come)
followed by the key (with its NUL) and the value; totsize = keysize +
len(value) + 16, so the entry ends 8 + totsize bytes after its header
starts.  Readers look entries up by key and skip unknown ones.  Values:
fundamental types little-endian as in memory, vectors and strings as u32
count + elements, Eigen matrices as u32 rows + u32 cols + column-major data,
maps as u32 count + (key, value) pairs, the dictionary as b"Dict" + the
vector of words.  Readers also accept the untagged version-0 encodings of
the model block (vocabWeights, alpha, alphas, eta, K) and of documents
(b"Docu", weight, words, wOrder, then Zs, wordWeights).
"""
import io
import struct

import numpy as np

TAG_MAGIC = 0x4B545054  # "TPTK"
TAG_VERSION = 0x00010001
TW_KEYS = {"one": b"one\0\0", "idf": b"idf\0\0", "pmi": b"pmi\0\0"}
MODEL_KEY = b"LDA\0\0"


class FormatError(ValueError):
    """The data is not a model file of this layout (tomoto's UnfitException)."""


# ------------------------------------------------------------------ writer --
class _Out:
    def __init__(self):
        self.parts = []

    def raw(self, b):
        self.parts.append(bytes(b))

    def u16(self, x):
        self.parts.append(struct.pack("<H", x))

    def u32(self, x):
        self.parts.append(struct.pack("<I", x))

    def u64(self, x):
        self.parts.append(struct.pack("<Q", x))

    def f32(self, x):
        self.parts.append(struct.pack("<f", x))

    def vec(self, a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        self.u32(len(a))
        self.parts.append(a.tobytes())

    def string(self, s):
        b = s.encode("utf-8") if isinstance(s, str) else bytes(s)
        self.u32(len(b))
        self.parts.append(b)

    def matrix(self, a, rows, cols, dtype):
        """Eigen matrix: rows, cols, column-major data (`a` given in that order)."""
        self.u32(rows)
        self.u32(cols)
        self.parts.append(np.ascontiguousarray(a, dtype=dtype).tobytes())

    def bytes(self):
        return b"".join(self.parts)


def _tagged(out, entries):
    """entries: list of (name, value bytes) written as one tagged list."""
    n = len(entries)
    for i, (name, value) in enumerate(entries):
        key = name.encode("ascii") + b"\0"
        out.raw(struct.pack("<IIQII", TAG_MAGIC, TAG_VERSION, len(key) + len(value) + 16, len(key), n - 1 - i))
        out.raw(key)
        out.raw(value)


def _val(fn, *args):
    o = _Out()
    fn(o, *args)
    return o.bytes()


def dump_model(st):
    """Bytes of a model file.  `st` holds: tw ('one' | 'idf' | 'pmi'), vocab
    (list of str, prepared order), vocab_cf, vocab_df, real_v, global_step,
    extra (bytes), vocab_weights, alpha (the scalar alpha), alphas, eta, K,
    eta_by_word (dict str -> list of float), burn_in, optim_interval, nk (K),
    nkw (real_v x K), docs (None, or a list of dicts with weight, words,
    w_order, raw_str, orig_word_pos, orig_word_len, uid, zs, word_weights)."""
    tw = st["tw"]
    cnt = np.int32 if tw == "one" else np.float32
    K, V = int(st["K"]), int(st["real_v"])
    out = _Out()
    out.raw(MODEL_KEY)
    out.raw(TW_KEYS[tw])

    def dict_val(o):
        o.raw(b"Dict")
        o.u32(len(st["vocab"]))
        for w in st["vocab"]:
            o.string(w)

    _tagged(out, [
        ("dict", _val(dict_val)),
        ("vocabCf", _val(lambda o: o.vec(st["vocab_cf"], np.uint64))),
        ("vocabDf", _val(lambda o: o.vec(st["vocab_df"], np.uint64))),
        ("realV", _val(lambda o: o.u64(V))),
        ("globalStep", _val(lambda o: o.u64(int(st["global_step"])))),
        ("extra", _val(lambda o: o.vec(np.frombuffer(st["extra"], np.uint8), np.uint8))),
    ])

    def eta_by_word(o):
        o.u32(len(st["eta_by_word"]))
        for w, pr in st["eta_by_word"].items():
            o.string(w)
            o.vec(pr, np.float32)

    alphas = np.asarray(st["alphas"], np.float32)
    _tagged(out, [
        ("vocabWeights", _val(lambda o: o.vec(st["vocab_weights"], np.float32))),
        ("alpha", _val(lambda o: o.f32(float(st["alpha"])))),
        ("alphas", _val(lambda o: o.matrix(alphas, len(alphas), 1, np.float32))),
        ("eta", _val(lambda o: o.f32(float(st["eta"])))),
        ("K", _val(lambda o: o.u16(K))),
        ("etaByWord", _val(eta_by_word)),
        ("burnIn", _val(lambda o: o.u32(int(st["burn_in"])))),
        ("optimInterval", _val(lambda o: o.u32(int(st["optim_interval"])))),
    ])
    out.matrix(np.asarray(st["nk"]), K, 1, cnt)
    nkw = np.asarray(st["nkw"]).reshape(V, K)
    out.matrix(nkw, K, V, cnt)  # word-major rows are Eigen's column-major K x V
    docs = st.get("docs")
    if docs is None:
        out.u32(0)
    else:
        out.u32(len(docs))
        for d in docs:
            _tagged(out, [
                ("weight", _val(lambda o: o.f32(float(d.get("weight", 1.0))))),
                ("words", _val(lambda o: o.vec(d["words"], np.uint32))),
                ("wOrder", _val(lambda o: o.vec(d["w_order"], np.uint32))),
                ("rawStr", _val(lambda o: o.string(d.get("raw_str", b"")))),
                ("origWordPos", _val(lambda o: o.vec(d.get("orig_word_pos", ()), np.uint32))),
                ("origWordLen", _val(lambda o: o.vec(d.get("orig_word_len", ()), np.uint16))),
                ("docUid", _val(lambda o: o.string(d.get("uid", "")))),
            ])
            _tagged(out, [
                ("Zs", _val(lambda o: o.vec(d["zs"], np.uint16))),
                ("wordWeights", _val(lambda o: o.vec(d.get("word_weights", ()), np.float32))),
            ])
    return out.bytes()


# ------------------------------------------------------------------ reader --
class _In:
    def __init__(self, data):
        self.b = memoryview(data)
        self.pos = 0

    def take(self, n):
        if n < 0 or self.pos + n > len(self.b):
            raise FormatError("truncated model data at byte %d" % self.pos)
        v = self.b[self.pos:self.pos + n]
        self.pos += n
        return v

    def peek(self, n):
        return bytes(self.b[self.pos:self.pos + n])

    def u16(self):
        return struct.unpack("<H", self.take(2))[0]

    def u32(self):
        return struct.unpack("<I", self.take(4))[0]

    def u64(self):
        return struct.unpack("<Q", self.take(8))[0]

    def f32(self):
        return struct.unpack("<f", self.take(4))[0]

    def vec(self, dtype):
        n = self.u32()
        it = np.dtype(dtype).itemsize
        return np.frombuffer(self.take(n * it), dtype=dtype).copy()

    def string(self):
        n = self.u32()
        return bytes(self.take(n))

    def matrix(self, dtype):
        rows, cols = self.u32(), self.u32()
        it = np.dtype(dtype).itemsize
        return rows, cols, np.frombuffer(self.take(rows * cols * it), dtype=dtype).copy()

    def key(self, k):
        got = bytes(self.take(len(k)))
        if got != k:
            raise FormatError("%r is needed but %r" % (k, got))


def _read_tagged(r):
    """readTaggedDataMap: {key: _In positioned on the value}; leaves r at the
    end of the last entry."""
    out = {}
    while True:
        start = r.pos
        if len(r.b) - start < 24:
            raise FormatError("truncated tagged data")
        magic, version, totsize, keysize, trailing = struct.unpack("<IIQII", r.take(24))
        if magic != TAG_MAGIC:
            r.pos = start
            raise FormatError("tagged data key is not found")
        key = bytes(r.take(keysize))
        end = start + 8 + totsize
        if end > len(r.b) or totsize < keysize + 16:
            raise FormatError("tagged entry %r overruns the data" % key)
        sub = _In(r.b[:end])
        sub.pos = r.pos
        out[key.rstrip(b"\0").decode("ascii", "replace")] = sub
        r.pos = end
        if trailing == 0:
            return out


def _is_tagged(r):
    return len(r.b) - r.pos >= 4 and struct.unpack("<I", r.peek(4))[0] == TAG_MAGIC


def parse_model(data):
    """Inverse of dump_model; returns the same fields (docs None when the
    file holds no documents)."""
    r = _In(data)
    r.key(MODEL_KEY)
    twk = r.peek(5)
    tw = {v: k for k, v in TW_KEYS.items()}.get(twk)
    if tw is None:
        raise FormatError("unknown term weighting key %r" % twk)
    r.take(5)
    if not _is_tagged(r):
        raise FormatError("model files written before tomotopy 0.6.0 are not supported")
    t = _read_tagged(r)
    st = {"tw": tw}
    if "dict" in t:
        d = t["dict"]
        d.key(b"Dict")
        st["vocab"] = [d.string().decode("utf-8") for _ in range(d.u32())]
    else:
        st["vocab"] = []
    st["vocab_cf"] = t["vocabCf"].vec(np.uint64) if "vocabCf" in t else np.zeros(0, np.uint64)
    st["vocab_df"] = t["vocabDf"].vec(np.uint64) if "vocabDf" in t else np.zeros(0, np.uint64)
    st["real_v"] = t["realV"].u64() if "realV" in t else 0
    st["global_step"] = t["globalStep"].u64() if "globalStep" in t else 0
    st["extra"] = bytes(t["extra"].vec(np.uint8)) if "extra" in t else b""
    # model block: tagged version 1, or the version-0 sequence
    st["eta_by_word"] = {}
    st["burn_in"], st["optim_interval"] = 0, 10
    if _is_tagged(r):
        m = _read_tagged(r)
        st["vocab_weights"] = m["vocabWeights"].vec(np.float32) if "vocabWeights" in m else np.zeros(0, np.float32)
        st["alpha"] = m["alpha"].f32()
        _, cols, st["alphas"] = m["alphas"].matrix(np.float32)
        if cols != 1:
            raise FormatError("alphas must be a column vector")
        st["eta"] = m["eta"].f32()
        st["K"] = m["K"].u16()
        if "etaByWord" in m:
            e = m["etaByWord"]
            for _ in range(e.u32()):
                w = e.string().decode("utf-8")
                st["eta_by_word"][w] = e.vec(np.float32)
        if "burnIn" in m:
            st["burn_in"] = m["burnIn"].u32()
        if "optimInterval" in m:
            st["optim_interval"] = m["optimInterval"].u32()
    else:
        st["vocab_weights"] = r.vec(np.float32)
        st["alpha"] = r.f32()
        _, cols, st["alphas"] = r.matrix(np.float32)
        st["eta"] = r.f32()
        st["K"] = r.u16()
    cnt = np.int32 if tw == "one" else np.float32
    K, V = st["K"], st["real_v"]
    rows, cols, nk = r.matrix(cnt)
    if (rows, cols) != (K, 1):
        raise FormatError("n_k has shape %dx%d, expected %dx1" % (rows, cols, K))
    rows, cols, nkw = r.matrix(cnt)
    if (rows, cols) != (K, V):
        raise FormatError("n_kw has shape %dx%d, expected %dx%d" % (rows, cols, K, V))
    st["nk"], st["nkw"] = nk, nkw.reshape(V, K)
    n_docs = r.u32()
    docs = []
    for _ in range(n_docs):
        d = {}
        if _is_tagged(r):
            b = _read_tagged(r)
            d["weight"] = b["weight"].f32() if "weight" in b else 1.0
            d["words"] = b["words"].vec(np.uint32)
            d["w_order"] = b["wOrder"].vec(np.uint32) if "wOrder" in b else np.zeros(0, np.uint32)
            d["raw_str"] = b["rawStr"].string() if "rawStr" in b else b""
            d["orig_word_pos"] = b["origWordPos"].vec(np.uint32) if "origWordPos" in b else np.zeros(0, np.uint32)
            d["orig_word_len"] = b["origWordLen"].vec(np.uint16) if "origWordLen" in b else np.zeros(0, np.uint16)
            d["uid"] = b["docUid"].string().decode("utf-8") if "docUid" in b else ""
            l = _read_tagged(r)
            d["zs"] = l["Zs"].vec(np.uint16)
            d["word_weights"] = l["wordWeights"].vec(np.float32) if "wordWeights" in l else np.zeros(0, np.float32)
        else:
            r.key(b"Docu")
            d["weight"] = r.f32()
            d["words"] = r.vec(np.uint32)
            d["w_order"] = r.vec(np.uint32)
            d["raw_str"], d["uid"] = b"", ""
            d["orig_word_pos"], d["orig_word_len"] = np.zeros(0, np.uint32), np.zeros(0, np.uint16)
            d["zs"] = r.vec(np.uint16)
            d["word_weights"] = r.vec(np.float32)
        if len(d["zs"]) != len(d["words"]):
            raise FormatError("a document has %d words but %d topics" % (len(d["words"]), len(d["zs"])))
        docs.append(d)
    st["docs"] = docs if n_docs else None
    return st


def write_file(path, st):
    with open(path, "wb") as f:
        f.write(dump_model(st))


def read_file(path):
    with open(path, "rb") as f:
        return parse_model(f.read())


def load_extra(extra):
    """init_params from the `extra` block (tomotopy pickles a dict there).
    Only plain built-in values are accepted; anything else yields {}."""
    import pickle

    class Restricted(pickle.Unpickler):
        def find_class(self, module, name):
            if module == "builtins" and name in ("dict", "list", "tuple", "set", "frozenset", "str", "int", "float",
                                                  "bool", "complex", "bytes"):
                import builtins
                return getattr(builtins, name)
            raise pickle.UnpicklingError("%s.%s is not a plain value" % (module, name))

    if not extra:
        return {}
    try:
        v = Restricted(io.BytesIO(extra)).load()
    except Exception:
        return {}
    return v if isinstance(v, dict) else {}
