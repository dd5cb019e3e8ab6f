"""Document and Corpus types of the LDAModel surface (tomotopy/utils.py:13-190).

A Document is a view of one training document of a model, or a document made
by ``LDAModel.make_doc`` for fold-in.  Topic read-outs follow tomoto's
_getTopicsByDoc (LDAModel.hpp:1145-1159) and the Document accessors of
src/python/handler/py_utils.cpp:952-1011.
"""
import warnings

import numpy as np

_NO_TOPIC_MSG = ("This document has no topic information. Call `infer()` method passing this document as an "
                 "argument first!")


class Document:
    """One document: word ids in original order, topics, and its model."""

    def __init__(self, model, word_ids, doc_index=None, uid=None):
        self._model = model
        self._word_ids = np.asarray(word_ids, dtype=np.int64)
        self._doc_index = doc_index  # index into the model's training documents, or None
        self._theta = None           # fold-in result (unnormalised n_dk + alpha is not kept)
        self._ndk = None
        self._z = None               # topics in original word order (int16, -1 = excluded)
        self._ll = None
        self.uid = uid or ""

    # -- identity -----------------------------------------------------------
    @property
    def words(self):
        """Vocabulary ids of the document's words, in the order given."""
        return self._word_ids.tolist()

    def __len__(self):
        return len(self._word_ids)

    def __getitem__(self, i):
        """The i-th word (str) of the document."""
        return self._model.vocabs[int(self._word_ids[i])]

    def __repr__(self):
        vocab = self._model.vocabs
        return "<tomotopy_amd.Document with words=\"%s\">" % " ".join(vocab[i] for i in self._word_ids[:20])

    @property
    def initialized(self):
        return self._doc_index is not None or self._theta is not None

    # -- topics ---------------------------------------------------------------
    @property
    def topics(self):
        """Topic of every word (int16; -1 for words outside the used vocabulary)."""
        if self._doc_index is not None:
            return self._model._doc_topics(self._doc_index)
        if self._z is None:
            warnings.warn(_NO_TOPIC_MSG, RuntimeWarning, stacklevel=2)
            return np.full(len(self._word_ids), -1, dtype=np.int16)
        return self._z

    def get_topic_dist(self, normalize=True):
        """(n_dk + alpha) / (L_d + sum(alpha)) (LDAModel.hpp:1145-1159)."""
        if self._doc_index is not None:
            return self._model._ctx.doc_topic_dist(self._doc_index, normalize).tolist()
        elif self._ndk is None:
            warnings.warn(_NO_TOPIC_MSG, RuntimeWarning, stacklevel=2)
            return [0.0] * self._model.k
        else:
            ndk = self._ndk.astype(np.float32)
        alpha = np.asarray(self._model.alpha, dtype=np.float32)
        num = ndk + alpha
        if not normalize:
            return num.tolist()
        asum = np.float32(0.0)
        for a in alpha:
            asum = np.float32(asum + a)
        return (num / (np.float32(ndk.sum()) + asum)).astype(np.float32).tolist()

    def get_topics(self, top_n=10):
        dist = self.get_topic_dist()
        if not any(dist):
            return []
        order = np.argsort(-np.asarray(dist), kind="stable")[:top_n]
        return [(int(k), float(dist[k])) for k in order]

    def get_ll(self):
        """Document part of the log-likelihood (getLLDocs, LDAModel.hpp:616-631), FP64."""
        if self._doc_index is None:
            return float(self._ll) if self._ll is not None else float("nan")
        return self._model._doc_ll(self._doc_index)

    def get_words(self, top_n=10):
        ids, cnt = np.unique(self._word_ids, return_counts=True)
        order = np.argsort(-cnt, kind="stable")[:top_n]
        vocab = self._model.vocabs
        return [(vocab[ids[i]], float(cnt[i]) / len(self._word_ids)) for i in order]


class RawDocument:
    """One document of a Corpus: its words (str) and uid (tomotopy.utils.Document
    for a corpus that is not bound to a model)."""

    def __init__(self, words, uid=""):
        self.words = list(words)
        self.uid = uid or ""

    def __len__(self):
        return len(self.words)

    def __iter__(self):
        return iter(self.words)

    def __getitem__(self, i):
        return self.words[i]

    def __repr__(self):
        return "<tomotopy_amd.utils.RawDocument with words=\"%s\">" % " ".join(self.words[:20])


class DocumentList(list):
    """A model's documents: indexed by position, by slice, or by uid."""

    def __init__(self, docs, uid_map=None):
        super().__init__(docs)
        self._uid_map = uid_map or {}

    def __getitem__(self, key):
        if isinstance(key, str):
            if key not in self._uid_map:
                raise KeyError("cannot find a document with uid = {}".format(repr(key)))
            return list.__getitem__(self, self._uid_map[key])
        if isinstance(key, slice):
            return DocumentList(list.__getitem__(self, key))
        return list.__getitem__(self, key)


class ModelCorpus:
    """The documents a model took from a Corpus (the result of add_corpus):
    each item is the model's Document (topics and log-likelihood after
    training); by position, slice or iteration.  It can be added to a model
    again with add_corpus."""

    def __init__(self, model, indices):
        self._model = model
        self._idx = list(indices)

    def __len__(self):
        return len(self._idx)

    def __iter__(self):
        return (self._model._document(i) for i in self._idx)

    def __getitem__(self, i):
        if isinstance(i, slice):
            return ModelCorpus(self._model, self._idx[i])
        return self._model._document(self._idx[i])


class Corpus:
    """A list of raw documents (words and uid) for LDAModel(corpus=...),
    add_corpus and infer (tomotopy/utils.py Corpus for pre-tokenised words)."""

    def __init__(self, docs=None):
        self._docs = []
        self._uids = []
        for d in docs or []:
            self.add_doc(d)

    def add_doc(self, words, uid=None):
        if isinstance(words, str):
            raise ValueError("`words` should be an iterable of str.")
        self._docs.append(list(words))
        self._uids.append(uid or "")
        return len(self._docs) - 1

    def __len__(self):
        return len(self._docs)

    def __iter__(self):
        return (RawDocument(w, u) for w, u in zip(self._docs, self._uids))

    def __getitem__(self, i):
        if isinstance(i, slice):
            sub = Corpus()
            sub._docs = self._docs[i]
            sub._uids = self._uids[i]
            return sub
        return RawDocument(self._docs[i], self._uids[i])


class InferredCorpus(list):
    """Documents made and inferred from a Corpus (infer(corpus) result)."""
