"""Spatial smoothing of the initial pressure (mirrors kwave/utils/filters.py:549-615)."""
import numpy as np

from .signals import get_win


def smooth(a, restore_max=False, window_type="Blackman"):
    """Multiply the spectrum of ``a`` by a rotated window (symmetric iff the axis length is odd).

    With ``restore_max`` the result is rescaled so that max|a_sm| == max|a|.
    """
    a = np.asarray(a)
    if a.dtype == bool:
        a = a.astype(int)
    if not np.all(np.isfinite(a)):
        raise ValueError("smooth input must be finite")
    orig = a.shape
    sq = np.squeeze(a)
    shape = sq.shape
    sym = np.array([s % 2 == 1 for s in shape])
    n_arg = shape[0] if len(shape) == 1 else shape
    sym_arg = bool(sym[0]) if len(shape) == 1 else sym
    win, _ = get_win(n_arg, type_=window_type, rotation=sq.ndim > 1, symmetric=sym_arg)
    win = np.abs(win).reshape(shape)
    out = np.real(np.fft.ifftn(np.fft.fftn(sq) * np.fft.ifftshift(win))).reshape(orig)
    if restore_max:
        out = (np.abs(a).max() / np.abs(out).max()) * out
    return out
