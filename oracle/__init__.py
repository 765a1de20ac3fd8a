"""FP64 CPU oracle (test infrastructure; see oracle/sph_oracle.cpp)."""
