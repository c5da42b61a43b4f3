"""CPU oracle package (test infrastructure only; see gb_oracle.h)."""
