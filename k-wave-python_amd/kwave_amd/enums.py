"""Absorption-mode switch (mirrors kwave/enums.py:4-12)."""
from enum import Enum


class AlphaMode(str, Enum):
    NO_ABSORPTION = "no_absorption"
    NO_DISPERSION = "no_dispersion"
    STOKES = "stokes"

    def __str__(self):
        return self.value
