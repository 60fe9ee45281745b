"""CPU FP64 oracle of the reference searchlight Procrustes path (test infrastructure only)."""
