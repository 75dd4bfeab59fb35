"""CPU FP64 oracle for the subset IC-GN path (test infrastructure only)."""
