"""FP64 CPU oracle of the response path (test infrastructure only; see oracle.py)."""
