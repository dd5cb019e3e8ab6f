"""Host-side checks that need no GPU: the C-ABI library, corpus preparation,
the LDAModel surface's argument handling, and the synthetic workloads."""
import ctypes

import numpy as np
import pytest

from tomotopy_amd import _native, corpus, synth
from tomotopy_amd.models import LDAModel, ParallelScheme, TermWeight


def test_library_loads_and_exports_every_declared_symbol():
    lib = _native.load()
    names = _native.header_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
        assert n in _native._SIGS, "binding misses %s" % n
    assert lib.lda_version() == 1


def test_enums_match_tomotopy():
    assert [int(x) for x in TermWeight] == [0, 1, 2]
    assert (ParallelScheme.DEFAULT, ParallelScheme.NONE, ParallelScheme.COPY_MERGE, ParallelScheme.PARTITION) == \
        (0, 1, 2, 3)


def test_vocab_order_filters_and_rotation():
    cf = np.array([5, 1, 9, 9, 3, 0])
    df = np.array([2, 1, 3, 1, 3, 0])
    order, real_v = corpus.vocab_order(cf, df)
    # (9,3) > (9,1) > (5,2) > (3,3) > (1,1) > (0,0)
    assert order.tolist() == [2, 4, 0, 1, 3, 5] and real_v == 6
    order, real_v = corpus.vocab_order(cf, df, min_cf=2, rm_top=1)
    # kept by frequency: 2,3,0,4 ; filtered: 1,5 ; top-1 (word 2) rotated to the end
    new_to_old = np.argsort(order)
    assert new_to_old.tolist() == [3, 0, 4, 1, 5, 2] and real_v == 3


def test_build_csr_sorts_and_records_order():
    docs = [np.array([3, 1, 2, 1]), np.array([0])]
    words, off, w_orders = corpus.build_csr(docs, np.arange(4))
    assert words.tolist() == [1, 1, 2, 3, 0] and off.tolist() == [0, 4, 5]
    assert w_orders[0].tolist() == [3, 0, 2, 1]
    # vectorised path agrees
    w2, off2 = corpus.csr_from_arrays(np.array([3, 1, 2, 1, 0]), np.array([0, 4, 5], np.uint64), np.arange(4))
    assert w2.tolist() == words.tolist()


def test_prepare_arrays_matches_per_document_path():
    words, off, nw, _ = synth.workload("plain_k20", n_docs=30)
    ws, offs, real_v, order = corpus.prepare_arrays(words, off, nw)
    docs = [words[int(off[d]):int(off[d + 1])] for d in range(30)]
    cf, df = corpus.counts(docs, nw)
    order2, real_v2 = corpus.vocab_order(cf, df)
    w2, off2, _ = corpus.build_csr(docs, order2)
    assert real_v == real_v2 and np.array_equal(order, order2) and np.array_equal(ws, w2)


def test_synthetic_workload_shapes():
    words, off, nw, w = synth.workload("plain_k20")
    L = np.diff(off).astype(np.int64)
    assert len(L) == 1000 and L.min() >= 20 and L.max() <= 2000
    assert 150_000 < len(words) < 300_000 and nw <= 5000
    again = synth.workload("plain_k20")[0]
    assert np.array_equal(words, again)  # seeded


def test_model_argument_checks():
    with pytest.raises(ValueError):
        LDAModel(k=0)
    with pytest.raises(ValueError):
        LDAModel(k=3, alpha=[0.1, 0.2])
    with pytest.raises(ValueError):
        LDAModel(k=3, alpha=-1.0)
    with pytest.raises(ValueError):
        LDAModel(k=3, eta=0.0)
    with pytest.raises(ValueError):
        LDAModel(k=3, seed="x")
    assert LDAModel(tw=TermWeight.PMI, k=3).tw == int(TermWeight.PMI)
    assert LDAModel(tw="idf", k=3).tw == int(TermWeight.IDF)
    with pytest.raises(ValueError):
        LDAModel(tw="bm25", k=3)
    m = LDAModel(k=3, seed=1)
    with pytest.raises(ValueError):
        m.add_doc("a string")
    with pytest.raises(ValueError):
        m.add_doc([[0, 1]])  # tomotopy issue 187: non-str words
    assert m.add_doc([]) is None
    with pytest.raises(ValueError):
        m.add_doc([], ignore_empty_words=False)
    assert m.add_doc(["a", "b", "a"]) == 0
    with pytest.raises(RuntimeError):
        m.make_doc(["a"])
    with pytest.raises(RuntimeError):
        m.infer([])
    assert m.ll_per_word == 0.0 and m.num_words == 0


def test_device_calls_fail_loudly_without_gpu():
    if _native.device_count() > 0:
        pytest.skip("a HIP device is visible")
    with pytest.raises(RuntimeError):
        _native.Context(4, 10, 0.1, 0.01, 1)
    m = LDAModel(k=2, seed=1)
    m.add_doc(["a", "b"])
    with pytest.raises(RuntimeError):
        m.train(1)


def test_every_public_attribute_of_a_new_model_is_readable():
    """unit_test.py properties: every public attribute of a model that has not
    been trained can be read (no device is needed for that)."""
    import tomotopy_amd as tp
    mdl = tp.LDAModel(tw=0, min_df=2, rm_top=2, k=10)
    for attr in [a for a in dir(mdl) if not a.startswith("_")]:
        getattr(mdl, attr)


def test_uid_lookup_and_corpus_documents():
    """unit_test.py test_uid / test_empty_uid: documents keep their uid from
    a Corpus, the model's documents are found by uid, and a Corpus item
    carries its words and uid."""
    import pytest
    import tomotopy_amd as tp
    cps = tp.utils.Corpus()
    cps.add_doc("test text".split(), uid="001")
    cps.add_doc("test text".split(), uid="abc")
    cps.add_doc("test text".split(), uid="0x1f")
    mdl = tp.LDAModel(k=2, corpus=cps)
    assert len(cps) == len(mdl.docs)
    assert cps[0].uid == mdl.docs[0].uid == "001"
    assert mdl.docs["abc"].uid == "abc" and mdl.docs["0x1f"].uid == "0x1f"
    with pytest.raises(KeyError):
        mdl.docs["nope"]
    assert len(mdl.docs[:2]) == 2 and len(cps[:2]) == 2
    assert cps[1].words == ["test", "text"] and cps[1][0] == "test"
    empty = tp.utils.Corpus()
    for _ in range(3):
        empty.add_doc("test text".split())
    m2 = tp.LDAModel(corpus=empty)
    assert len(m2.docs) == 3 and empty[0].uid == m2.docs[0].uid == ""
    with pytest.raises(ValueError):
        mdl.add_doc("test text".split(), uid="001")  # a uid is unique in a model


def test_add_corpus_returns_the_model_documents():
    """unit_test.py test_empty_uid: add_corpus returns the model's documents,
    which can be added to a model again."""
    import tomotopy_amd as tp
    cps = tp.utils.Corpus()
    for _ in range(3):
        cps.add_doc("test text".split())
    mdl = tp.LDAModel()
    ccps = mdl.add_corpus(cps)
    assert len(ccps) == 3 and ccps[0].words == mdl.docs[0].words
    again = mdl.add_corpus(ccps)
    assert len(again) == 3 and len(mdl.docs) == 6
    assert [mdl.vocabs[i] for i in mdl.docs[5].words] == ["test", "text"]
    assert len(ccps[:2]) == 2
