"""tomotopy_amd: MI355X-native collapsed-Gibbs LDA behind tomotopy's LDAModel surface."""
__version__ = "0.1.0"
