"""CPU oracle of the LDA hot path -- test infrastructure only (see lda_oracle.cpp)."""
