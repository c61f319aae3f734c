"""CPU oracle package (test infrastructure only; see np_oracle.py)."""
