"""FP64 CPU oracle of the atmo preprocess (test infrastructure only)."""
