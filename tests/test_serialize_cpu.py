"""tomoto's model file layout on the CPU: byte-level structure of the
header, tagged lists and count matrices (restated from serializer.hpp and
the members TopicModel::_saveModel writes), round trips, tolerance of the
reader to reordered / unknown / missing tagged entries (lookup by key, as
readTaggedDataMap does), the version-0 encodings, and malformed input.
No model file written by the reference itself is available here, so byte
compatibility with tomotopy-written files is checked against the restated
layout only."""
import pickle
import struct

import numpy as np
import pytest

from tomotopy_amd import serialize as S


def _state(tw="one", K=4, V=6, n_docs=3, seed=0):
    rng = np.random.default_rng(seed)
    cnt = np.int32 if tw == "one" else np.float32
    docs = []
    for d in range(n_docs):
        n = int(rng.integers(1, 9))
        words = np.sort(rng.integers(0, V + 2, n)).astype(np.uint32)
        zs = rng.integers(0, K, n).astype(np.uint16)
        zs[words >= V] = 0xFFFF
        docs.append({"weight": 1.0, "words": words, "w_order": rng.permutation(n).astype(np.uint32),
                     "uid": "doc%d" % d if d % 2 else "", "zs": zs,
                     "word_weights": rng.random(n).astype(np.float32) if tw != "one" else np.zeros(0, np.float32)})
    return {"tw": tw, "vocab": ["w%d" % i for i in range(V + 2)] + ["été"],
            "vocab_cf": rng.integers(0, 100, V + 3).astype(np.uint64),
            "vocab_df": rng.integers(0, 10, V + 3).astype(np.uint64), "real_v": V, "global_step": 40,
            "extra": pickle.dumps({"k": K, "tw": 0}), "vocab_weights": rng.random(V).astype(np.float32)
            if tw != "one" else np.zeros(0, np.float32), "alpha": 0.1,
            "alphas": rng.random(K).astype(np.float32), "eta": 0.01, "K": K, "eta_by_word": {},
            "burn_in": 2, "optim_interval": 10, "nk": (rng.random(K) * 50).astype(cnt),
            "nkw": (rng.random((V, K)) * 9).astype(cnt), "docs": docs}


def _same(a, b):
    for k in a:
        if k in ("docs", "alpha", "eta"):  # floats: compared as float32 by the callers
            continue
        x, y = a[k], b[k]
        if isinstance(x, (str, bytes, list)):
            assert x == y, k
        elif isinstance(x, dict):
            assert set(x) == set(y), k
        else:
            assert np.array_equal(np.asarray(x, np.asarray(y).dtype), np.asarray(y)), k
    if a["docs"] is None:
        assert b["docs"] is None
        return
    assert len(a["docs"]) == len(b["docs"])
    for p, q in zip(a["docs"], b["docs"]):
        for k in ("words", "w_order", "zs", "word_weights"):
            assert np.array_equal(np.asarray(p[k]), q[k]), k
        assert p["uid"] == q["uid"]


@pytest.mark.parametrize("tw", ["one", "idf", "pmi"])
def test_round_trip(tw):
    st = _state(tw)
    st2 = S.parse_model(S.dump_model(st))
    _same(st, st2)
    assert st2["tw"] == tw
    assert st2["alpha"] == np.float32(0.1) and st2["eta"] == np.float32(0.01)
    st["docs"] = None  # full=False
    st3 = S.parse_model(S.dump_model(st))
    _same(st, st3)


def _entry(name, value, trailing):
    key = name.encode() + b"\0"
    return struct.pack("<IIQII", 0x4B545054, 0x00010001, len(key) + len(value) + 16, len(key), trailing) + key + value


def test_header_and_first_tagged_list_bytes():
    st = _state("one", K=2, V=1, n_docs=0)
    st["vocab"] = ["ab", "c"]
    st["vocab_cf"] = np.array([3, 1], np.uint64)
    st["vocab_df"] = np.array([2, 1], np.uint64)
    st["extra"] = b"xy"
    st["docs"] = None
    b = S.dump_model(st)
    assert b[:10] == b"LDA\0\0one\0\0"
    exp = (_entry("dict", b"Dict" + struct.pack("<I", 2) + struct.pack("<I", 2) + b"ab" + struct.pack("<I", 1) + b"c", 5)
           + _entry("vocabCf", struct.pack("<IQQ", 2, 3, 1), 4)
           + _entry("vocabDf", struct.pack("<IQQ", 2, 2, 1), 3)
           + _entry("realV", struct.pack("<Q", 1), 2)
           + _entry("globalStep", struct.pack("<Q", 40), 1)
           + _entry("extra", struct.pack("<I", 2) + b"xy", 0))
    assert b[10:10 + len(exp)] == exp
    # the "TPTK" magic reads as those four characters
    assert b[10:14] == b"TPTK"


def test_model_block_and_count_matrices_bytes():
    st = _state("one", K=3, V=2, n_docs=0)
    st["docs"] = None
    st["vocab_weights"] = np.zeros(0, np.float32)
    st["alphas"] = np.array([0.5, 0.25, 0.125], np.float32)
    st["nk"] = np.array([7, 8, 9], np.int32)
    st["nkw"] = np.array([[1, 2, 3], [4, 5, 6]], np.int32)  # word-major: word 0 has counts 1, 2, 3
    b = S.dump_model(st)
    model = (_entry("vocabWeights", struct.pack("<I", 0), 7)
             + _entry("alpha", struct.pack("<f", 0.1), 6)
             + _entry("alphas", struct.pack("<II3f", 3, 1, 0.5, 0.25, 0.125), 5)
             + _entry("eta", struct.pack("<f", 0.01), 4)
             + _entry("K", struct.pack("<H", 3), 3)
             + _entry("etaByWord", struct.pack("<I", 0), 2)
             + _entry("burnIn", struct.pack("<I", 2), 1)
             + _entry("optimInterval", struct.pack("<I", 10), 0))
    counts = struct.pack("<II3i", 3, 1, 7, 8, 9) + struct.pack("<II6i", 3, 2, 1, 2, 3, 4, 5, 6)
    tail = model + counts + struct.pack("<I", 0)
    assert b.endswith(tail)


def test_document_bytes():
    st = _state("pmi", K=2, V=3, n_docs=0)
    st["docs"] = [{"weight": 1.0, "words": np.array([0, 2], np.uint32), "w_order": np.array([1, 0], np.uint32),
                   "uid": "u", "zs": np.array([1, 0], np.uint16), "word_weights": np.array([0.5, 2.0], np.float32)}]
    b = S.dump_model(st)
    doc = (_entry("weight", struct.pack("<f", 1.0), 6)
           + _entry("words", struct.pack("<III", 2, 0, 2), 5)
           + _entry("wOrder", struct.pack("<III", 2, 1, 0), 4)
           + _entry("rawStr", struct.pack("<I", 0), 3)
           + _entry("origWordPos", struct.pack("<I", 0), 2)
           + _entry("origWordLen", struct.pack("<I", 0), 1)
           + _entry("docUid", struct.pack("<I", 1) + b"u", 0)
           + _entry("Zs", struct.pack("<IHH", 2, 1, 0), 1)
           + _entry("wordWeights", struct.pack("<Iff", 2, 0.5, 2.0), 0))
    assert b.endswith(struct.pack("<I", 1) + doc)


def test_reader_looks_tagged_entries_up_by_key():
    """Entries in another order, an unknown entry, and a missing optional one."""
    st = _state("one", K=2, V=1, n_docs=0)
    st["docs"] = None
    b = S.dump_model(st)
    # rebuild the model block in another order with an extra unknown key and without etaByWord
    start = b.index(b"vocabWeights\0") - 24
    end = b.index(b"optimInterval\0", start) + len(b"optimInterval\0") + 4  # the block's last entry
    entries = [("optimInterval", struct.pack("<I", 7)), ("future", b"\x01\x02\x03"),
               ("K", struct.pack("<H", 2)), ("alphas", struct.pack("<II2f", 2, 1, 0.5, 0.5)),
               ("eta", struct.pack("<f", 0.02)), ("alpha", struct.pack("<f", 0.5)),
               ("vocabWeights", struct.pack("<I", 0)), ("burnIn", struct.pack("<I", 1))]
    block = b"".join(_entry(n, v, len(entries) - 1 - i) for i, (n, v) in enumerate(entries))
    st2 = S.parse_model(b[:start] + block + b[end:])
    assert st2["optim_interval"] == 7 and st2["burn_in"] == 1 and st2["K"] == 2
    assert st2["eta"] == np.float32(0.02) and st2["eta_by_word"] == {}
    assert np.array_equal(st2["alphas"], np.array([0.5, 0.5], np.float32))


def test_version0_model_block_and_documents():
    st = _state("one", K=2, V=2, n_docs=0)
    st["docs"] = None
    b = S.dump_model(st)
    start = b.index(b"vocabWeights\0") - 24
    end = b.index(b"optimInterval\0", start) + len(b"optimInterval\0") + 4
    v0 = struct.pack("<I", 0) + struct.pack("<f", 0.3) + struct.pack("<II2f", 2, 1, 0.3, 0.3) + \
        struct.pack("<f", 0.05) + struct.pack("<H", 2)
    body = b[:start] + v0 + b[end:-4]
    doc0 = b"Docu" + struct.pack("<f", 1.0) + struct.pack("<III", 2, 0, 1) + struct.pack("<III", 2, 0, 1) + \
        struct.pack("<IHH", 2, 1, 1) + struct.pack("<I", 0)
    st2 = S.parse_model(body + struct.pack("<I", 1) + doc0)
    assert st2["alpha"] == np.float32(0.3) and st2["eta"] == np.float32(0.05)
    assert st2["optim_interval"] == 10 and st2["burn_in"] == 0
    d = st2["docs"][0]
    assert list(d["words"]) == [0, 1] and list(d["zs"]) == [1, 1] and d["uid"] == ""


def test_malformed_input():
    b = S.dump_model(_state("idf"))
    with pytest.raises(S.FormatError):
        S.parse_model(b[:-3])
    with pytest.raises(S.FormatError):
        S.parse_model(b"LDA\0\0xyz\0\0" + b[10:])
    with pytest.raises(S.FormatError):
        S.parse_model(b"HDP\0\0" + b[5:])
    with pytest.raises(S.FormatError):
        S.parse_model(b"")


def test_extra_unpickles_plain_values_only():
    assert S.load_extra(pickle.dumps({"k": 3, "alpha": [0.1, 0.2], "seed": None})) == \
        {"k": 3, "alpha": [0.1, 0.2], "seed": None}

    class Obj:
        pass
    import collections
    assert S.load_extra(pickle.dumps({"x": collections.OrderedDict()})) == {}
    assert S.load_extra(b"") == {}
    assert S.load_extra(b"not a pickle") == {}
