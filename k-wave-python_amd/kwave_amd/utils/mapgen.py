"""Geometry helpers for solver inputs (mirrors kwave/utils/mapgen.py:505-719,822-955)."""
import numpy as np

from ..data import Vector


def _pixel_dim(n, origin_size="single", shift=1):
    """Pixel coordinates along one axis; kwave/utils/mapgen.py:909-955."""
    if n % 2 == 0:
        if origin_size == "single":
            return np.arange(-n / 2, n / 2) if shift == 1 else np.arange(-n / 2 + 1, n / 2 + 1)
        return np.hstack([np.arange(-n / 2 + 1, 1), np.arange(0, n / 2)])
    if origin_size == "single":
        return np.arange(-(n - 1) / 2, (n - 1) / 2 + 1)
    if shift == 1:
        return np.hstack([np.arange(-(n - 1) / 2, 1), np.arange(0, (n - 1) / 2)])
    return np.hstack([np.arange(-(n - 1) / 2 + 1, 1), np.arange(0, (n - 1) / 2 + 1)])


def make_pixel_map(grid_size, shift=None, origin_size="single"):
    """Distance of each pixel from the grid centre; kwave/utils/mapgen.py:822-907."""
    grid_size = [int(g) for g in grid_size]
    nd = len(grid_size)
    if nd not in (2, 3):
        raise ValueError("Grid size must be a 2 or 3 element vector.")
    shift = [1] * nd if shift is None else shift
    axes = [_pixel_dim(n, origin_size, s) for n, s in zip(grid_size, shift)]
    grids = np.meshgrid(*axes, indexing="ij")
    return np.sqrt(sum(g ** 2 for g in grids))


def make_disc(grid_size, center, radius, plot_disc=False):
    """Binary disc (wrapping at the edges); kwave/utils/mapgen.py:664-719."""
    gs = np.round(np.asarray(grid_size)).astype(int)
    c = np.round(np.asarray(center)).astype(int)
    if gs.size != 2 or c.size != 2:
        raise ValueError("grid_size and center must be 2D")
    c = np.array([c[i] if c[i] != 0 else gs[i] // 2 + 1 for i in range(2)])
    if not (np.all(c > 0) and np.all(c <= gs)):
        raise ValueError("Disc center must be within grid.")
    disc = np.zeros(gs, dtype=bool)
    disc[make_pixel_map(gs, shift=[0, 0]) <= radius] = True
    return np.roll(disc, tuple(c - np.ceil(gs / 2).astype(int)), axis=(0, 1))


def make_ball(grid_size, ball_center, radius, plot_ball=False, binary=False):
    """Ball in a 3-D grid (wrapping at the edges); kwave/utils/mapgen.py:505-554."""
    gs = np.asarray(grid_size).astype(int)
    c = np.asarray(ball_center).astype(int).copy()
    for i in range(3):
        if c[i] == 0:
            c[i] = gs[i] // 2 + 1
    ball = np.zeros(gs, dtype=bool if binary else int)
    ball[make_pixel_map(gs, shift=[0, 0, 0]) <= radius] = 1
    return np.roll(ball, tuple(c - np.ceil(gs / 2).astype(int)), axis=(0, 1, 2))


def make_cart_circle(radius, num_points, center_pos=Vector([0, 0]), arc_angle=2 * np.pi, plot_circle=False):
    """2 x num_points Cartesian points on a circle or arc; kwave/utils/mapgen.py:611-661."""
    full = arc_angle == 2 * np.pi
    n_steps = num_points if full else num_points - 1
    ang = np.arange(0, num_points) * arc_angle / n_steps + np.pi / 2
    circ = np.concatenate([radius * np.cos(ang[None, :]), radius * np.sin(-ang[None, :])])
    circ = circ + np.asarray(center_pos, dtype=float)[:, None]
    return np.squeeze(circ)
