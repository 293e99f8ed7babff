"""CPU oracle package (test infrastructure only; see snac_oracle.h)."""
