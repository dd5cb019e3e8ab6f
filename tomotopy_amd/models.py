"""LDAModel with tomotopy's Python surface (tomotopy/models.py:48-658).

Host-side responsibilities (dictionary, documents, vocabulary preparation,
argument checks, callbacks, read-outs) live here; every sampling sweep, the
alpha fixed point, the log-likelihood and fold-in run on the GPU through the
C ABI (include/tomotopy_amd.h) of the HIP library.  There is no CPU sampler:
without the library or a HIP device, train/infer raise.
"""
import math
import os
import pickle
import sys
import warnings
from copy import deepcopy
from enum import IntEnum

import numpy as np

from . import _native
from . import corpus as _corpus
from . import serialize as _serialize
from .utils import Corpus, Document, DocumentList, InferredCorpus, ModelCorpus

__version__ = "0.1.0"


class TermWeight(IntEnum):
    """tomotopy.TermWeight (tomotopy/__init__.py:10-34)."""
    ONE = 0
    IDF = 1
    PMI = 2


class ParallelScheme(IntEnum):
    """tomotopy.ParallelScheme (tomotopy/__init__.py:36-64).

    On the device: NONE (or workers=1) runs one worker; PARTITION/DEFAULT with
    workers=W >= 2 runs tomoto's partition schedule with W device workers;
    COPY_MERGE with workers=W >= 2 tomoto's copy_merge with W workers;
    PARTITION or COPY_MERGE with workers=0 that scheme with the device's own
    worker count; DEFAULT with workers=0 the device-wide schedule (every
    workgroup a worker with a private n_k, exact atomics on n_kw, at most an
    eighth of the documents in flight), which tomotopy does not have.
    """
    DEFAULT = 0
    NONE = 1
    COPY_MERGE = 2
    PARTITION = 3


def _convert_term_weight(tw):
    if isinstance(tw, str):
        try:
            return getattr(TermWeight, tw.upper())
        except AttributeError:
            raise ValueError(f"Unknown term weight: {tw}")
    return TermWeight(int(tw))


def _format_numpy(arr, prefix=""):
    arr = np.array(arr)
    return ("\n" + prefix).join(str(arr).split("\n"))


class LDAModel:
    """Latent Dirichlet Allocation by collapsed Gibbs sampling on an MI355X.

The first train() with workers=0 and parallel=DEFAULT warns once per process
that it runs the device-wide schedule rather than tomotopy's partition.

> * Blei, D.M., Ng, A.Y., & Jordan, M.I. (2003). Latent dirichlet allocation. JMLR 3, 993-1022.
> * Newman, D., Asuncion, A., Smyth, P., & Welling, M. (2009). Distributed algorithms for topic models. JMLR 10, 1801-1828.
"""

    _dw_warned = False

    def __init__(self, tw="one", min_cf=0, min_df=0, rm_top=0, k=1, alpha=0.1, eta=0.01, seed=None, corpus=None,
                 transform=None):
        '''Parameters
----------
tw : Union[int, tomotopy_amd.TermWeight]
    term weighting scheme (TermWeight.ONE, IDF or PMI)
min_cf : int
    minimum collection frequency of words
min_df : int
    minimum document frequency of words
rm_top : int
    the number of top words to be removed
k : int
    the number of topics between 1 ~ 32767
alpha : Union[float, Iterable[float]]
    hyperparameter of Dirichlet distribution for document-topic
eta : float
    hyperparameter of Dirichlet distribution for topic-word
seed : int
    random seed
corpus : tomotopy_amd.utils.Corpus
    a list of documents to be added into the model
transform : Callable[dict, dict]
    a callable object to manipulate arbitrary keyword arguments
'''
        self.init_params = deepcopy({k_: v for k_, v in locals().items() if k_ not in ("self", "corpus", "transform")
                                     and not k_.startswith("_")})
        self.init_params["version"] = __version__
        tw = _convert_term_weight(tw)
        # LDAModel ctor checks (LDAModel.hpp:964-984), k range (models.py:84)
        if not isinstance(k, (int, np.integer)) or not (1 <= int(k) <= _native.LDA_MAX_K):
            raise ValueError("wrong K value (K = %r); the device sampler supports 1 ~ %d" % (k, _native.LDA_MAX_K))
        k = int(k)
        if isinstance(alpha, (int, float, np.floating, np.integer)):
            alphas = np.full(k, alpha, dtype=np.float32)
        else:
            try:
                alphas = np.asarray(list(alpha), dtype=np.float32)
            except (TypeError, ValueError):
                raise ValueError("`alpha` must be an instance of `float` or `List[float]` with length `k` (given %r)"
                                 % (alpha,))
            if len(alphas) != k:
                raise ValueError("`alpha` must be an instance of `float` or `List[float]` with length `k` (given %r)"
                                 % (alpha,))
        if np.any(alphas <= 0):
            raise ValueError("wrong alpha value")
        if not float(eta) > 0:
            raise ValueError("wrong eta value (eta = %f)" % eta)
        if seed is not None and not isinstance(seed, (int, np.integer)):
            raise ValueError("`seed` must be an integer or None.")
        self._seed_given = seed is not None
        self._seed = int(seed) if seed is not None else int.from_bytes(os.urandom(8), "little")
        self._tw = tw
        self._k = k
        self._alphas = alphas
        self._alpha0 = float(alphas[0])  # tomoto's scalar `alpha` (LDAModel.hpp:965)
        self._vocab_weights = np.zeros(0, np.float32)
        self._eta = float(np.float32(eta))
        self._min_cf, self._min_df, self._rm_top = int(min_cf), int(min_df), int(rm_top)
        self._optim_interval = 10
        self._burn_in = 0
        self._global_step = 0
        # dictionary and raw documents (ids in order of first appearance)
        self._dict = {}
        self._vocab = []
        self._raw_docs = []
        self._uids = []
        self._uid_map = {}
        self._prepared = False
        self._ctx = None
        self._real_v = 0
        self._weights = None
        self._eta_by_word = {}  # word -> K priors (setWordPrior keeps the first one set)
        self._cache = {}
        if corpus is not None:
            self.add_corpus(corpus, transform)

    # ------------------------------------------------------------ documents --
    def _word_id(self, w, add):
        i = self._dict.get(w)
        if i is None and add:
            i = len(self._vocab)
            self._dict[w] = i
            self._vocab.append(w)
        return i

    def add_doc(self, words, ignore_empty_words=True, uid=None):
        '''Add a new document; returns its index, or None for an ignored empty document.'''
        if self._prepared:
            raise RuntimeError("cannot add_doc() after train()")
        if isinstance(words, str):
            raise ValueError("`words` should be an iterable of str.")
        try:
            ws = list(words)
        except TypeError:
            raise ValueError("`words` should be an iterable of str.")
        for w in ws:
            if not isinstance(w, str):
                raise ValueError("`words` should be an iterable of str.")
        if not ws:
            if ignore_empty_words:
                return None
            raise ValueError("Either `words` or `raw` must be filled.")
        if uid:
            if uid in self._uid_map:
                raise ValueError("there is a document with uid = '%s' already." % uid)
            self._uid_map[uid] = len(self._raw_docs)
        ids = np.fromiter((self._word_id(w, True) for w in ws), dtype=np.int64, count=len(ws))
        self._raw_docs.append(ids)
        self._uids.append(uid or "")
        return len(self._raw_docs) - 1

    def add_corpus(self, corpus, transform=None):
        """Add every document of a Corpus (or of another add_corpus result);
        returns the model's new documents as a ModelCorpus."""
        if self._prepared:
            raise RuntimeError("cannot add_corpus() after train()")
        if isinstance(corpus, ModelCorpus):
            src = [([corpus._model._vocab[i] for i in d._word_ids], d.uid) for d in corpus]
        elif isinstance(corpus, Corpus):
            src = list(zip(corpus._docs, corpus._uids))
        else:
            raise ValueError("`corpus` must be an instance of `tomotopy_amd.utils.Corpus`")
        idx = []
        for words, uid in src:
            r = self.add_doc(words, uid=uid or None)
            if r is not None:
                idx.append(r)
        return ModelCorpus(self, idx)

    def _document(self, i):
        """Training document i (its topics once the model is prepared)."""
        if self._prepared:
            return Document(self, self._raw_docs[i], doc_index=i, uid=self._uids[i])
        return Document(self, self._raw_docs[i], uid=self._uids[i])

    def make_doc(self, words):
        '''Return a Document for an unseen document (out-of-dictionary words are dropped).'''
        if not self._prepared:
            raise RuntimeError("`train()` should be called before `make_doc()`.")
        if isinstance(words, str):
            raise ValueError("`words` should be an iterable of str.")
        ids = [self._dict[w] for w in words if w in self._dict]
        new_ids = self._order[np.asarray(ids, dtype=np.int64)] if ids else np.zeros(0, np.int64)
        return Document(self, new_ids)

    # -------------------------------------------------------------- prepare --
    def _prepare(self):
        """tomoto prepare(): vocabulary order, token CSR, initial topics on the device."""
        n_words = len(self._vocab)
        cf, df = _corpus.counts(self._raw_docs, n_words)
        order, real_v = _corpus.vocab_order(cf, df, self._min_cf, self._min_df, self._rm_top)
        if real_v == 0:
            print("[warn] No valid vocabs in the model!", file=sys.stderr)
        self._order = order
        inv = np.empty(n_words, dtype=np.int64)
        inv[order] = np.arange(n_words)
        self._vocab = [self._vocab[i] for i in inv]
        self._dict = {w: i for i, w in enumerate(self._vocab)}
        self._cf = cf[inv]
        self._df = df[inv]
        self._raw_docs = [order[d] for d in self._raw_docs]
        self._order = np.arange(n_words)  # ids are now the prepared ones
        words, off, w_orders = _corpus.build_csr(self._raw_docs, np.arange(n_words))
        self._words, self._off, self._w_orders = words, off, w_orders
        self._real_v = int(real_v)
        self._n_valid = int(np.count_nonzero(words < real_v))
        self._weights = None
        self._vocab_weights = _corpus.vocab_weights(words, off, self._real_v, self._tw.name.lower())
        if self._tw != TermWeight.ONE:
            self._weights = _corpus.term_weights(words, off, self._real_v, self._tw.name.lower())
        self._ctx = _native.Context(self._k, self._real_v, self._alphas, self._eta, self._seed, tw=int(self._tw))
        self._upload_priors(self._ctx)
        self._ctx.upload_corpus(words, off, weights=self._weights)
        self._ctx.set_optim(self._optim_interval, self._burn_in)
        self._ctx.global_step = self._global_step
        self._prepared = True
        self._cache = {}

    def _upload_priors(self, ctx):
        """prepareWordPriors (LDAModel.hpp:724-737): the prior columns of the
        words that survived preparation (ids < realV; the others are skipped)."""
        if not self._eta_by_word:
            return
        words = [w for w in self._eta_by_word if w in self._dict]
        ids = np.array([self._dict[w] for w in words], dtype=np.uint32)
        pri = np.array([self._eta_by_word[w] for w in words], dtype=np.float32).reshape(len(ids), self._k)
        ctx.set_word_priors(ids, pri)

    def set_word_prior(self, word, prior):
        '''Set the topic prior of `word` (a list of K floats, or a dict {topic: prior} with eta elsewhere);
call before `train` (py_LDA.cpp:241-269, LDAModel.hpp:1031-1045).'''
        if self._prepared:
            raise RuntimeError("cannot set_word_prior() after train()")
        if isinstance(prior, dict):
            pr = [self._eta] * self._k
            for k, v in prior.items():
                if not isinstance(k, (int, np.integer)) or not 0 <= int(k) < self._k:
                    raise ValueError("`prior` must be a dict of {topic_id: float}")
                try:
                    pr[int(k)] = float(v)
                except (TypeError, ValueError):
                    raise ValueError("`prior` must be a dict of {topic_id: float}")
        else:
            try:
                pr = [float(x) for x in prior]
            except (TypeError, ValueError):
                raise ValueError("`prior` must be a list of floats with len = k")
        if len(pr) != self._k:
            raise ValueError("priors.size() must be equal to K.")
        if any(not (x >= 0) for x in pr):
            raise ValueError("priors must not be less than 0.")
        if not isinstance(word, str):
            raise ValueError("`word` must be str")
        self._word_id(word, True)
        self._eta_by_word.setdefault(word, [float(np.float32(x)) for x in pr])

    def get_word_prior(self, word):
        '''The topic prior of `word` (LDAModel.hpp:1047-1062): after `train` the word's eta column
(its prior, or eta everywhere) when any prior is set; an empty list otherwise.'''
        if self._prepared and self._eta_by_word:
            i = self._dict.get(word)
            if i is None or i >= self._real_v:
                return []
            pr = self._eta_by_word.get(word)
            return list(pr) if pr is not None else [self._eta] * self._k
        pr = self._eta_by_word.get(word)
        return list(pr) if pr is not None else []

    def _ensure_prepared(self):
        if not self._prepared:
            self._prepare()

    # ---------------------------------------------------------------- train --
    def train(self, iterations=10, workers=0, parallel=0, freeze_topics=False, callback_interval=10, callback=None,
              show_progress=False):
        '''Train the model with `iterations` Gibbs sweeps on the GPU (py_LDA.cpp:277-320).
With workers=0 and parallel=DEFAULT the sweeps run the device-wide schedule,
which tomotopy does not have (its DEFAULT is PARTITION): pass
parallel=PARTITION for tomoto's schedule with the device's worker count.'''
        if workers == 0 and int(parallel) == int(ParallelScheme.DEFAULT) and not LDAModel._dw_warned:
            LDAModel._dw_warned = True
            warnings.warn("workers=0 with ParallelScheme.DEFAULT runs the device-wide schedule (shared n_kw with exact "
                          "atomics, private n_k per workgroup), not tomotopy's partition schedule; pass "
                          "parallel=ParallelScheme.PARTITION for that", RuntimeWarning, stacklevel=2)
        if self._seed_given and workers != 1:
            warnings.warn("The training result may differ even with fixed seed if `workers` != 1.", RuntimeWarning,
                          stacklevel=2)
        if show_progress:
            user_cb = callback
            if user_cb is None:
                callback = LDAModel._show_progress
            else:
                def callback(*args):
                    user_cb(*args)
                    LDAModel._show_progress(*args)
        self._ensure_prepared()
        if callback is not None and not callable(callback):
            raise ValueError("`callback` should be a callable object")
        if callback is None or callback_interval <= 0:
            callback_interval = iterations
        parallel = int(parallel)
        it = 0
        while it < iterations:
            if callback is not None:
                callback(self, it, iterations)
            n = min(callback_interval, iterations - it)
            try:
                self._ctx.train(n, workers=int(workers), scheme=parallel)
            finally:
                # sweeps before a failure changed the state
                self._cache = {}
                self._global_step = self._ctx.global_step
                self._alphas = self._ctx.get_alpha()
            it += n
        if callback is not None:
            callback(self, iterations, iterations)
        self._global_step = self._ctx.global_step
        self._alphas = self._ctx.get_alpha()

    def _show_progress(self, current_iteration, total_iteration):
        print("Iteration %d/%d  LLPW: %.6f" % (current_iteration, total_iteration, self.ll_per_word), file=sys.stderr)

    # --------------------------------------------------------------- state ---
    def _state(self, what):
        if what not in self._cache:
            if self._tw != TermWeight.ONE:
                z, ndk, nkw, nk = self._ctx.download_state_f()
            else:
                z, ndk, nkw, nk = self._ctx.download_state(z=what == "z", ndk=what == "ndk", nkw=what == "nkw",
                                                           nk=what == "nk")
            self._cache[what] = {"z": z, "ndk": ndk, "nkw": nkw, "nk": nk}[what]
        return self._cache[what]

    def _doc_topics(self, d):
        b, e = int(self._off[d]), int(self._off[d + 1])
        z = self._state("z")[b:e].astype(np.int32)
        words = self._words[b:e]
        z = np.where(words < self._real_v, z, -1).astype(np.int16)
        return z[self._w_orders[d]]

    def _doc_ndk(self, d):
        return self._state("ndk")[d]

    def _doc_ll(self, d):
        from scipy.special import gammaln  # FP64 closed form of getLLDocs for one document
        a = np.asarray(self._alphas, np.float64)
        ndk = self._doc_ndk(d).astype(np.float64)
        n = ndk.sum()
        if self._weights is not None:  # getSumWordWeight: sequential float sum of the weights
            b, e = int(self._off[d]), int(self._off[d + 1])
            acc = np.float32(0.0)
            for w, x in zip(self._words[b:e], self._weights[b:e]):
                if w < self._real_v:
                    acc = np.float32(acc + x)
            n = float(acc)
        return float(gammaln(a.sum()) - gammaln(n + a.sum()) + np.sum(gammaln(ndk + a) - gammaln(a)))

    # ------------------------------------------------------------ read-outs --
    def get_topic_word_dist(self, topic_id, normalize=True):
        '''(n_kw + eta) / (n_k + V*eta) for every used word (LDAModel.hpp:828-840), read out on the device.'''
        self._ensure_prepared()
        if not 0 <= topic_id < self._k:
            raise ValueError("topic_id must < K")
        return self._ctx.topic_word_dist(int(topic_id), normalize).tolist()

    def get_topic_words(self, topic_id, top_n=10, return_id=False):
        '''The top_n words of a topic by get_topic_word_dist (TopicModel.hpp:285-288, 777-815); words
of equal probability are listed by ascending id (tomoto's std::sort leaves their order unspecified).'''
        dist = np.asarray(self.get_topic_word_dist(topic_id))
        order = np.argsort(-dist, kind="stable")[:top_n]
        if return_id:
            return [(self._vocab[i], int(i), float(dist[i])) for i in order]
        return [(self._vocab[i], float(dist[i])) for i in order]

    def get_count_by_topics(self):
        '''Tokens assigned to each topic (LDAModel.hpp:815-826), counted on the device.'''
        self._ensure_prepared()
        return self._ctx.count_by_topic().astype(np.int64).tolist()

    # ---------------------------------------------------------------- infer --
    def infer(self, doc, iterations=100, tolerance=-1, workers=0, parallel=0, together=False, transform=None):
        '''Topic distributions of unseen documents (py_LDA.cpp:345-411, LDAModel.hpp:843-954).'''
        if not self._prepared:
            raise RuntimeError("cannot infer with untrained model")
        kw = {"together": bool(together), "workers": int(workers), "parallel": int(parallel)}
        if isinstance(doc, Corpus):
            docs = [self.make_doc(d) for d in doc]
            _, ll = self._infer_docs(docs, iterations, **kw)
            return InferredCorpus(docs), ll
        if isinstance(doc, Document):
            if doc._model is not self:
                raise ValueError("`doc` was from another model, not fit to this model")
            if doc._doc_index is not None:
                return doc.get_topic_dist(), None
            dists, ll = self._infer_docs([doc], iterations, **kw)
            return dists[0], ll[0]
        try:
            docs = list(doc)
        except TypeError:
            raise ValueError("`doc` must be tomotopy_amd.Document type or list of tomotopy_amd.Document")
        for d in docs:
            if not isinstance(d, Document):
                raise ValueError("`doc` must be tomotopy_amd.Document type or list of tomotopy_amd.Document")
            if d._model is not self:
                raise ValueError("`doc` was from another model, not fit to this model")
        dists, ll = self._infer_docs(docs, iterations, **kw)
        return (dists, ll[0]) if together else (dists, ll)

    def _infer_docs(self, docs, iterations, mode=0, together=False, workers=0, parallel=0):
        """Fold-in on the device.  together=False: every document against the
        frozen model (one batched launch); together=True: the documents sampled
        jointly on a copy of the model, one log-likelihood for all of them
        (LDAModel.hpp:852-892)."""
        if not docs:
            return [], ([0.0] if together else [])
        lens = [len(d._word_ids) for d in docs]
        off = np.zeros(len(docs) + 1, dtype=np.uint64)
        off[1:] = np.cumsum(lens)
        words = np.empty(int(off[-1]), dtype=np.uint32)
        perms = []
        for i, d in enumerate(docs):
            perm = np.argsort(d._word_ids, kind="stable")
            words[int(off[i]):int(off[i + 1])] = d._word_ids[perm]
            perms.append(perm)
        wts = None
        if self._tw != TermWeight.ONE:
            # initializeDocState (LDAModel.hpp:794-809): weights from the model's vocabulary weights
            wts = _corpus.term_weights(words, off, self._real_v, self._tw.name.lower(), vocab_w=self._vocab_weights)
        if together:
            theta, ll_all, z = self._ctx.infer_together(words, off, iterations, infer_seed=0, workers=workers,
                                                        scheme=parallel, weights=wts)
            ll = np.array([ll_all])
        else:
            theta, ll, z = self._ctx.infer(words, off, iterations, infer_seed=0, mode=mode, want_z=True, weights=wts)
        out = []
        for i, d in enumerate(docs):
            b, e = int(off[i]), int(off[i + 1])
            zz = z[b:e].astype(np.int32)
            zz = np.where(words[b:e] < self._real_v, zz, -1)
            orig = np.empty(len(zz), dtype=np.int16)
            orig[perms[i]] = zz
            d._z = orig
            if wts is not None:
                d._ndk = np.bincount(zz[zz >= 0], weights=wts[b:e][zz >= 0], minlength=self._k).astype(np.float32)
            else:
                d._ndk = np.bincount(zz[zz >= 0], minlength=self._k)
            d._theta = theta[i]
            d._ll = ll[0] if together else ll[i]
            out.append(theta[i].tolist())
        return out, ll.tolist()

    # ------------------------------------------------------------ properties --
    @property
    def k(self):
        return self._k

    @property
    def alpha(self):
        return [float(a) for a in self._alphas]

    @property
    def eta(self):
        return self._eta

    @property
    def tw(self):
        return int(self._tw)

    @property
    def global_step(self):
        return self._global_step

    @property
    def optim_interval(self):
        return self._optim_interval

    @optim_interval.setter
    def optim_interval(self, value):
        if int(value) < 0 or int(value) > 0x7FFFFFFF:
            raise ValueError("wrong value")
        self._optim_interval = int(value)
        if self._ctx is not None:
            self._ctx.set_optim(self._optim_interval, self._burn_in)

    @property
    def burn_in(self):
        return self._burn_in

    @burn_in.setter
    def burn_in(self, value):
        if int(value) < 0 or int(value) > 0x7FFFFFFF:
            raise ValueError("wrong value")
        self._burn_in = int(value)
        if self._ctx is not None:
            self._ctx.set_optim(self._optim_interval, self._burn_in)

    @property
    def ll_per_word(self):
        '''getLL() / weightedN (TopicModel.hpp:757-760), evaluated in FP64 on the device.'''
        if not self._prepared or self._n_valid == 0:
            return 0.0
        a, b = self._ctx.loglik()
        return (a + b) / self._ctx.weighted_n()

    @property
    def perplexity(self):
        return math.exp(-self.ll_per_word)

    @property
    def num_words(self):
        return self._n_valid if self._prepared else 0

    @property
    def num_vocabs(self):
        return self._real_v if self._prepared else 0

    @property
    def docs(self):
        """The training documents: by position, slice, or uid (docs["uid"])."""
        if not self._prepared:
            return DocumentList([Document(self, d, uid=u) for d, u in zip(self._raw_docs, self._uids)], self._uid_map)
        return DocumentList([Document(self, self._raw_docs[i], doc_index=i, uid=self._uids[i])
                             for i in range(len(self._raw_docs))], self._uid_map)

    @property
    def vocabs(self):
        return list(self._vocab)

    @property
    def used_vocabs(self):
        return list(self._vocab[:self._real_v]) if self._prepared else []

    @property
    def vocab_freq(self):
        return self._cf.tolist() if self._prepared else _corpus.counts(self._raw_docs, len(self._vocab))[0].tolist()

    @property
    def vocab_df(self):
        return self._df.tolist() if self._prepared else _corpus.counts(self._raw_docs, len(self._vocab))[1].tolist()

    @property
    def used_vocab_freq(self):
        return np.asarray(self._cf[:self._real_v]) if self._prepared else np.zeros(0, np.int64)

    @property
    def used_vocab_df(self):
        return np.asarray(self._df[:self._real_v]) if self._prepared else np.zeros(0, np.int64)

    @property
    def used_vocab_weighted_freq(self):
        '''Sum of the token weights of every used word (TopicModel.hpp:914-926).'''
        if not self._prepared:
            return np.zeros(0, np.float64)
        if self._weights is None:
            return self.used_vocab_freq.astype(np.float64)
        valid = self._words < self._real_v
        out = np.zeros(self._real_v, np.float64)
        np.add.at(out, self._words[valid].astype(np.int64), self._weights[valid].astype(np.float64))
        return out

    @property
    def removed_top_words(self):
        if not self._prepared or not self._rm_top:
            return []
        return list(self._vocab[len(self._vocab) - min(self._rm_top, len(self._vocab)):])

    # ------------------------------------------------------------ save/load --
    def _plain_init_params(self):
        def plain(v):
            if isinstance(v, IntEnum):
                return int(v)
            if isinstance(v, np.generic):
                return v.item()
            if isinstance(v, np.ndarray):
                return v.tolist()
            if isinstance(v, (list, tuple)):
                return type(v)(plain(x) for x in v)
            if isinstance(v, dict):
                return {k: plain(x) for k, x in v.items()}
            return v
        return plain(getattr(self, "init_params", {}))

    def _file_state(self, full):
        self._ensure_prepared()
        tw = self._tw.name.lower()
        st = {
            "tw": tw, "vocab": list(self._vocab), "vocab_cf": np.asarray(self._cf, np.uint64),
            "vocab_df": np.asarray(self._df, np.uint64), "real_v": self._real_v,
            "global_step": self.global_step, "extra": pickle.dumps(self._plain_init_params()),
            "vocab_weights": self._vocab_weights, "alpha": self._alpha0,
            "alphas": np.asarray(self._alphas, np.float32), "eta": self._eta, "K": self._k,
            "eta_by_word": dict(self._eta_by_word),
            "burn_in": self._burn_in, "optim_interval": self._optim_interval,
            "nk": self._state("nk"), "nkw": self._state("nkw"), "docs": None,
        }
        if full:
            z = self._state("z").copy()
            z[self._words >= self._real_v] = 0xFFFF  # tomoto's non_topic_id for removed words
            docs = []
            for d in range(len(self._off) - 1):
                b, e = int(self._off[d]), int(self._off[d + 1])
                docs.append({"weight": 1.0, "words": self._words[b:e], "w_order": self._w_orders[d],
                             "uid": self._uids[d] or "", "zs": z[b:e],
                             "word_weights": self._weights[b:e] if self._weights is not None else ()})
            st["docs"] = docs
        return st

    def save(self, filename, full=True):
        '''Save the model to `filename` in tomoto's model file layout (TopicModel.hpp:335-359).
If `full`, the documents and their topics are saved too (the model can be
trained further); otherwise only the topic parameters (inference only).'''
        _serialize.write_file(filename, self._file_state(full))

    def saves(self, full=True):
        '''The bytes `save` would write.'''
        return _serialize.dump_model(self._file_state(full))

    @classmethod
    def load(cls, filename):
        '''Load a model file written by `save` (or by tomotopy >= 0.6.0).'''
        with open(filename, "rb") as f:
            data = f.read()
        try:
            return cls._from_file_state(_serialize.parse_model(data))
        except _serialize.FormatError as e:
            raise RuntimeError("'%s' is not valid model file (%s)" % (filename, e)) from None

    @classmethod
    def loads(cls, data):
        '''Load a model from bytes produced by `saves`.'''
        try:
            return cls._from_file_state(_serialize.parse_model(bytes(data)))
        except _serialize.FormatError as e:
            raise RuntimeError("`data` is not valid model file (%s)" % e) from None

    @classmethod
    def _from_file_state(cls, st):
        """_loadModel + prepare(false) (TopicModel.hpp:361-395, 943-947): counts
        n_kw and n_k come from the file, n_dk is rebuilt from the topics."""
        params = _serialize.load_extra(st["extra"])
        seed = params.get("seed")
        K, V = int(st["K"]), int(st["real_v"])
        mdl = cls(tw=st["tw"], k=K, alpha=[float(a) for a in st["alphas"]] if K > 1 else float(st["alphas"][0]),
                  eta=float(st["eta"]), seed=seed if isinstance(seed, int) else None)
        mdl.init_params = params
        mdl._min_cf = int(params.get("min_cf", 0) or 0)
        mdl._min_df = int(params.get("min_df", 0) or 0)
        mdl._rm_top = int(params.get("rm_top", 0) or 0)
        mdl._alpha0 = float(st["alpha"])
        mdl._vocab = list(st["vocab"])
        mdl._dict = {w: i for i, w in enumerate(mdl._vocab)}
        n_words = len(mdl._vocab)
        mdl._cf = np.zeros(n_words, np.int64)
        mdl._df = np.zeros(n_words, np.int64)
        mdl._cf[:len(st["vocab_cf"])] = np.asarray(st["vocab_cf"][:n_words], np.int64)
        mdl._df[:len(st["vocab_df"])] = np.asarray(st["vocab_df"][:n_words], np.int64)
        mdl._real_v = V
        mdl._order = np.arange(n_words)
        mdl._vocab_weights = np.asarray(st["vocab_weights"], np.float32)
        mdl._global_step = int(st["global_step"])
        mdl._burn_in = int(st["burn_in"])
        mdl._optim_interval = int(st["optim_interval"])
        docs = st["docs"] or []
        lens = [len(d["words"]) for d in docs]
        off = np.zeros(len(docs) + 1, np.uint64)
        off[1:] = np.cumsum(lens)
        cat = lambda key, dt: (np.concatenate([np.asarray(d[key], dt) for d in docs]) if docs else np.zeros(0, dt))
        words = cat("words", np.uint32)
        z = cat("zs", np.uint16)
        mdl._words, mdl._off = words, off
        mdl._w_orders = []
        mdl._raw_docs = []
        mdl._uids = []
        mdl._uid_map = {}
        for i, d in enumerate(docs):
            w = np.asarray(d["words"], np.int64)
            wo = np.asarray(d["w_order"], np.int64)
            if len(wo) != len(w):
                wo = np.arange(len(w))
            mdl._w_orders.append(wo)
            mdl._raw_docs.append(w[wo])
            mdl._uids.append(d["uid"])
            if d["uid"]:
                mdl._uid_map[d["uid"]] = i
        mdl._n_valid = int(np.count_nonzero(words < V))
        mdl._weights = None
        if mdl._tw != TermWeight.ONE:
            mdl._weights = cat("word_weights", np.float32) if docs else np.zeros(0, np.float32)
            if len(mdl._weights) != len(words):
                raise _serialize.FormatError("term-weighted documents without word weights")
        z = np.where(words < V, z, 0).astype(np.uint16)
        mdl._eta_by_word = {w: [float(x) for x in pr] for w, pr in st["eta_by_word"].items()}
        mdl._ctx = _native.Context(K, V, mdl._alphas, mdl._eta, mdl._seed, tw=int(mdl._tw))
        mdl._upload_priors(mdl._ctx)
        mdl._ctx.upload_corpus(words, off, z, weights=mdl._weights)
        if mdl._tw == TermWeight.ONE:
            mdl._ctx.set_counts(st["nkw"], st["nk"])
        else:
            mdl._ctx.set_counts_f(st["nkw"], st["nk"])
        mdl._ctx.set_optim(mdl._optim_interval, mdl._burn_in)
        mdl._ctx.global_step = mdl._global_step
        mdl._prepared = True
        mdl._cache = {}
        return mdl

    # --------------------------------------------------------------- misc ----
    def copy(self):
        '''A deep copy with its own device state.'''
        new = LDAModel.__new__(LDAModel)
        for key, val in self.__dict__.items():
            if key in ("_ctx", "_cache"):
                continue
            new.__dict__[key] = deepcopy(val)
        new._ctx = None
        new._cache = {}
        if self._prepared:
            # the device state is copied as is (topics, n_dk, n_kw, n_k, the
            # float counts of IDF / PMI models, priors, alpha, global step),
            # as tomoto's copy constructor copies globalState and the documents
            new._ctx = self._ctx.copy()
        return new

    def summary(self, initial_hp=True, params=True, topic_word_top_n=5, file=None, flush=False):
        print("<Basic Info>", file=file)
        print("| LDAModel (current version: %s)" % __version__, file=file)
        print("| %d docs, %d words" % (len(self._raw_docs), self.num_words), file=file)
        print("| Total Vocabs: %d, Used Vocabs: %d" % (len(self._vocab), len(self.used_vocabs)), file=file)
        if self._prepared and self._real_v:
            p = self.used_vocab_freq / max(1, self.used_vocab_freq.sum())
            print("| Entropy of words: %.5f" % float(-(p * np.log(p + 1e-20)).sum()), file=file)
        print("| Removed Vocabs: %s" % (" ".join(self.removed_top_words) or "<NA>"), file=file)
        print("|", file=file)
        print("<Training Info>", file=file)
        print("| Iterations: %d, Burn-in steps: %d" % (self.global_step, self.burn_in), file=file)
        print("| Optimization Interval: %d" % self.optim_interval, file=file)
        print("| Log-likelihood per word: %.5f" % self.ll_per_word, file=file)
        print("|", file=file)
        if initial_hp:
            print("<Initial Parameters>", file=file)
            for key, v in self.init_params.items():
                print("| %s: %s" % (key, v), file=file)
            print("|", file=file)
        if params:
            print("<Parameters>", file=file)
            print("| alpha (Dirichlet prior on the per-document topic distributions)\n|  %s"
                  % _format_numpy(self.alpha, "|  "), file=file)
            print("| eta (Dirichlet prior on the per-topic word distribution)\n|  %.5g" % self.eta, file=file)
            print("|", file=file)
        if topic_word_top_n > 0 and self._prepared:
            print("<Topics>", file=file)
            cnt = self.get_count_by_topics()
            for k in range(self.k):
                words = " ".join(w for w, _ in self.get_topic_words(k, top_n=topic_word_top_n))
                print("| #%d (%d) : %s" % (k, cnt[k], words), file=file)
            print("|", file=file)
        print(file=file, flush=flush)
