"""Unit conversion (mirrors kwave/utils/conversion.py:19-35)."""
import math

import numpy as np


def db2neper(alpha, y=1):
    """dB/(MHz^y cm) -> Np/((rad/s)^y m)."""
    return 100.0 * np.asarray(alpha, dtype=float) * (1e-6 / (2.0 * math.pi)) ** y / (20.0 * np.log10(np.exp(1)))


def neper2db(alpha, y=1):
    """Inverse of db2neper."""
    return 20.0 * np.log10(np.exp(1)) * np.asarray(alpha, dtype=float) * (2.0 * math.pi * 1e6) ** y / 100.0
