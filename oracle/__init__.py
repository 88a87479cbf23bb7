"""CPU oracle (test infrastructure only; see qcbc_oracle.c)."""
