"""ORACLE — test infrastructure only (see gip_oracle.h)."""
