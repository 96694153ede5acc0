"""CPU oracle of the matrix-free path (test infrastructure; see oracle/src/oracle.hpp)."""
