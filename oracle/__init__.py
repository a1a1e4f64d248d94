"""CPU oracle (test infrastructure): see sbe_oracle.h."""
