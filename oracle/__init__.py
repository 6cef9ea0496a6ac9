"""CPU oracle (test infrastructure only; see lfm_oracle.py's header)."""
