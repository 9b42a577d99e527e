"""FP64 CPU oracle for Improved CEEMDAN (test infrastructure; see oracle.py)."""
