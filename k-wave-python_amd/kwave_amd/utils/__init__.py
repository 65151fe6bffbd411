"""Host-side helpers used to build solver inputs (FP64 NumPy)."""
