"""CPU oracle of the qvm pure-state path (test infrastructure only; see qvm_oracle.c)."""
