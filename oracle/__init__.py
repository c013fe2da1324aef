"""CPU oracle of the reionization path. TEST INFRASTRUCTURE ONLY (see oracle.h)."""
