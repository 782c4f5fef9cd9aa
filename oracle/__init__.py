"""CPU oracle (test infrastructure only): see bp_oracle.py."""
