"""Reconstruction callers of the solver."""
from .time_reversal import TimeReversal

__all__ = ["TimeReversal"]
