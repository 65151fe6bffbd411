"""Unit conversion (mirrors kwave/utils/conversion.py:19-35)."""
import math

import numpy as np


def db2neper(alpha, y=1):
    """dB/(MHz^y cm) -> Np/((rad/s)^y m)."""
    return 100.0 * np.asarray(alpha, dtype=float) * (1e-6 / (2.0 * math.pi)) ** y / (20.0 * np.log10(np.exp(1)))


def neper2db(alpha, y=1):
    """Inverse of db2neper."""
    return 20.0 * np.log10(np.exp(1)) * np.asarray(alpha, dtype=float) * (2.0 * math.pi * 1e6) ** y / 100.0


def cart2grid(kgrid, cart_data, axisymmetric=False, *, order="C"):
    """Nearest grid point of each Cartesian point (kwave/utils/conversion.py:165-330).

    Positions are measured from the grid centre (index N//2 on every axis), so a
    point at coordinate x lands on index round(x / d) + N // 2.  Returns
    (grid_data, order_index, reorder_index): grid_data holds the 1-based point
    number at each hit grid point (0 elsewhere); order_index lists the point
    numbers in the flat order of the hit points ("C" or "F" flattening); and
    reorder_index (n, 1) gives, for each hit point in that order, its point number.
    Only order="C" and order="F" are accepted; axisymmetric grids are not.
    """
    if axisymmetric:
        raise NotImplementedError("axisymmetric grids are not part of this package")
    if order not in ("C", "F"):
        raise ValueError("order must be 'C' or 'F'")
    cart = np.atleast_2d(np.asarray(cart_data, dtype=float))
    shape = tuple(int(n) for n in np.atleast_1d(kgrid.N))
    spacing = [float(d) for d in np.atleast_1d(kgrid.spacing)]
    if cart.shape[0] != len(shape):
        raise ValueError(f"Cartesian points must have {len(shape)} rows for a {len(shape)}-D grid")
    idx = [np.round(cart[a] / spacing[a]).astype(np.int64) + shape[a] // 2 for a in range(len(shape))]
    for a, ia in enumerate(idx):
        if ia.size and (ia.min() < 0 or ia.max() >= shape[a]):
            raise ValueError("Cartesian points must lie within the grid defined by kgrid.")
    grid = np.zeros(shape)
    grid[tuple(idx)] = np.arange(1, cart.shape[1] + 1)
    flat = grid.ravel(order=order)
    hit = flat[flat != 0]
    return grid, hit.astype(np.int64), hit.reshape(-1, 1)
