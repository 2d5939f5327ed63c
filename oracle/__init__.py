"""CPU oracle (test infrastructure only; see cbs_oracle.py)."""
