"""FP64 CPU oracle (test infrastructure only; see pw_oracle.py header)."""
