"""ORACLE — test infrastructure only: the CPU restatement of the reference renderer (see oracle.py)."""
