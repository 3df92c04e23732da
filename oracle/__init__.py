"""FP64 CPU oracle of the reference warp path (test infrastructure only)."""
