"""Small vector type with .x/.y/.z accessors (mirrors kwave/data.py:7-52)."""
import numpy as np


class Vector(np.ndarray):
    """1- to 3-element numeric vector; kwave/data.py:7-52."""

    def __new__(cls, elements):
        elements = list(elements)
        if not 1 <= len(elements) <= 3:
            raise ValueError("Vector holds 1 to 3 elements")
        return np.array(elements).view(cls)

    def _get(self, i):
        return self[i] if self.shape[0] > i else np.nan

    def _set(self, i, v):
        if self.shape[0] > i:
            self[i] = v

    x = property(lambda s: s[0], lambda s, v: s.__setitem__(0, v))
    y = property(lambda s: s._get(1), lambda s, v: s._set(1, v))
    z = property(lambda s: s._get(2), lambda s, v: s._set(2, v))

    def append(self, value):
        return Vector(list(self) + [value])
