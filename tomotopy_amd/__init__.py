"""tomotopy_amd: MI355X-native collapsed-Gibbs LDA behind tomotopy's LDAModel surface.

    import tomotopy_amd as tp
    mdl = tp.LDAModel(k=20)
    for line in open("corpus.txt"):
        mdl.add_doc(line.split())
    mdl.train(100)
    print(mdl.ll_per_word, mdl.get_topic_words(0))
"""
from .models import LDAModel, ParallelScheme, TermWeight, __version__
from .utils import Corpus, Document

__all__ = ["LDAModel", "ParallelScheme", "TermWeight", "Corpus", "Document", "__version__"]
