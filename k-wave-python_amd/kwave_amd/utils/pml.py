"""Perfectly matched layer profiles (mirrors kwave/utils/pml.py:6-159)."""
import numpy as np

from .math import largest_prime_factor


def get_pml(Nx, dx, dt, c, pml_size, pml_alpha, staggered, dimension=1, axisymmetric=False):
    """1-D PML multiplier exp(-alpha (c/dx) s(x)^4 dt/2) at both ends, 1 inside.

    Returns shape (Nx, 1) for dimension 1, (1, Nx) for 2 and (1, 1, Nx) for 3,
    like the reference.
    """
    Nx, L = int(Nx), int(pml_size)
    x = np.arange(1, L + 1, dtype=float)
    pos = x + 0.5 if staggered else x
    left = pml_alpha * (c / dx) * (((pos - L - 1) / (0 - L)) ** 4)
    right = pml_alpha * (c / dx) * ((pos / L) ** 4)
    prof = np.ones((1, Nx))
    if not axisymmetric:
        prof[:, :L] = np.exp(-left * dt / 2)
    prof[:, Nx - L:] = np.exp(-right * dt / 2)
    if dimension == 1:
        return prof.T
    if dimension == 3:
        return prof.reshape(1, 1, Nx)
    return prof


def get_optimal_pml_size(grid_size, pml_range=None, axisymmetric=None):
    """PML size in ``pml_range`` giving the smallest largest prime factor of N + 2*pml."""
    from ..kgrid import kWaveGrid

    if isinstance(grid_size, kWaveGrid):
        grid_size = grid_size.N
    grid_size = list(np.atleast_1d(grid_size))
    nd = len(grid_size)
    if not 1 <= nd <= 3:
        raise ValueError("Grid dimensions must be given as a 1, 2, or 3 element vector.")
    rng = np.round([10, 40] if pml_range is None else pml_range).astype(int)
    if len(rng) != 2 or rng[0] < 0 or rng[1] <= rng[0]:
        raise ValueError("pml_range must be an increasing pair of non-negative values")
    sizes = np.arange(rng[0], rng[1] + 1)
    facs = np.zeros((nd, sizes.size))
    for d in range(nd):
        for j, s in enumerate(sizes):
            if isinstance(axisymmetric, str) and d == 2:
                n = (grid_size[d] + s) * 4 if axisymmetric == "WSWA" else (grid_size[d] + s) * 2 - 2
            else:
                n = grid_size[d] + 2 * s
            facs[d, j] = largest_prime_factor(int(n))
    return sizes[np.argmin(facs, 1)]
