"""CPU FP64 oracle of the VPFFT strain step (test infrastructure only)."""
