"""Computational grid (the part of kwave/kgrid.py the solver path reads).

Mirrors ``kWaveGrid.__init__`` (kwave/kgrid.py:31-90), ``setTime`` (:134-154),
``t_array`` (:92-132), ``makeDim`` (:404-430) and ``makeTime`` (:455-495).
"""
import math
import sys

import numpy as np

from .data import Vector


def _matlab_rem(x, y, rtol=1e-05, atol=1e-08):
    """MATLAB-style remainder with a tolerance for near-integer quotients (kwave/utils/matlab.py:7-33)."""
    if np.any(y == 0):
        return np.nan
    q = x / y
    if np.isclose(q, np.round(q), rtol=rtol, atol=atol).all():
        return np.zeros_like(x)
    return x - np.fix(q) * y


class kWaveGrid:
    CFL_DEFAULT = 0.3
    MACHINE_PRECISION = 100 * sys.float_info.epsilon

    def __init__(self, N, spacing):
        N, spacing = np.atleast_1d(N), np.atleast_1d(spacing)
        if not (N.ndim == 1 and spacing.ndim == 1 and 1 <= N.size <= 3 and N.size == spacing.size):
            raise ValueError("N and spacing must be 1-3 element vectors of equal length")
        self.N = N.astype(int)
        self.spacing = spacing.astype(float)
        self.dim = self.N.size
        self.nonuniform = False
        self.dt = "auto"
        self.Nt = "auto"

    # -- sizes ------------------------------------------------------------
    Nx = property(lambda s: s.N[0])
    Ny = property(lambda s: s.N[1] if s.N.size >= 2 else 0)
    Nz = property(lambda s: s.N[2] if s.N.size == 3 else 0)
    dx = property(lambda s: s.spacing[0])
    dy = property(lambda s: s.spacing[1] if s.spacing.size >= 2 else 0)
    dz = property(lambda s: s.spacing[2] if s.spacing.size == 3 else 0)

    @property
    def size(self):
        return Vector(self.N * self.spacing)

    @property
    def total_grid_points(self):
        return np.prod(self.N)

    @property
    def k_vec(self):
        return [self.makeDim(n, d) for n, d in zip(self.N, self.spacing)]

    # -- time -------------------------------------------------------------
    @property
    def t_array(self):
        if self.Nt == "auto" or self.dt == "auto":
            return "auto"
        return np.expand_dims(np.arange(0, self.Nt) * self.dt, axis=0)

    @t_array.setter
    def t_array(self, t_array):
        if isinstance(t_array, str) and t_array == "auto":
            self.Nt = "auto"
            self.dt = "auto"
            return
        t_array = np.asarray(t_array).ravel()
        dt = t_array[1] - t_array[0]
        if t_array[0] != 0:
            raise ValueError("t_array must begin at zero.")
        if (t_array[1:] - t_array[:-1] - dt).max() >= self.MACHINE_PRECISION:
            raise ValueError("t_array must be evenly spaced.")
        if dt <= 0:
            raise ValueError("t_array must be monotonically increasing.")
        self.Nt = t_array.size
        self.dt = dt

    def setTime(self, Nt, dt) -> None:
        if not (isinstance(Nt, (int, np.integer)) and Nt > 0):
            raise ValueError("Nt must be a positive integer.")
        if not dt > 0:
            raise ValueError("dt must be positive.")
        self.Nt = int(Nt)
        self.dt = dt

    @staticmethod
    def makeDim(num_points, spacing):
        """Wavenumber vector with DC at index N//2; kwave/kgrid.py:404-430."""
        n = int(num_points)
        if n % 2 == 0:
            nx = np.arange(-n / 2, n / 2) / n
        else:
            nx = np.arange(-(n - 1) / 2, (n - 1) / 2 + 1) / n
        nx[n // 2] = 0
        return ((2 * math.pi / spacing) * nx)[:, None]

    def makeTime(self, c, cfl=CFL_DEFAULT, t_end=None):
        """dt = cfl*min(d)/max(c); Nt = floor(t_end/dt)+1; kwave/kgrid.py:455-495."""
        c = np.array(c)
        c_min, c_max = np.min(c), np.max(c)
        if t_end is None:
            t_end = np.linalg.norm(self.size, ord=2) / c_min
        self.dt = cfl * np.min(self.spacing) / c_max
        self.Nt = int(np.floor(t_end / self.dt) + 1)
        if (np.floor(t_end / self.dt) != np.ceil(t_end / self.dt)) and (_matlab_rem(t_end, self.dt) == 0):
            self.Nt = self.Nt + 1
        return self.t_array, self.dt
