"""CPU oracle of the depth hot path (test infrastructure only)."""
