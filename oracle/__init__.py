"""CPU oracle of the projection hot path (test infrastructure only)."""
