"""CPU oracle (test infrastructure only): see amdat_oracle.c."""
