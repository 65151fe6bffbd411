"""Small math helpers (mirrors kwave/utils/math.py)."""
import numpy as np


def largest_prime_factor(n: int) -> int:
    """Largest prime factor by trial division; kwave/utils/math.py:121-138."""
    n = int(n)
    f = 2
    while f * f <= n:
        if n % f:
            f += 1
        else:
            n //= f
    return n


def gaussian(x, magnitude=None, mean=0, variance=1):
    """magnitude * exp(-(x-mean)^2 / (2 variance)); kwave/utils/math.py:348-380."""
    if magnitude is None:
        magnitude = (2 * np.pi * variance) ** -0.5
    return magnitude * np.exp(-((x - mean) ** 2) / (2 * variance))
