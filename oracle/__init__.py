"""CPU oracle of the hdfe-stream solve path (test infrastructure only; see
oracle.py)."""
