"""CPU oracle for the two-phase step (test infrastructure only; see oracle.py)."""
