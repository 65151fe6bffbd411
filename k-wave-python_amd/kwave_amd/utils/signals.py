"""Window and tone-burst generators (mirrors kwave/utils/signals.py:57-434)."""
import numpy as np
from scipy.interpolate import interp1d

from .math import gaussian

_COSINE_SERIES = {
    "Blackman-Harris": [0.35875, 0.48829, 0.14128, 0.01168],
    "Blackman-Nuttall": [0.3635819, 0.4891775, 0.1365995, 0.0106411],
    "Nuttall": [0.3635819, 0.4891775, 0.1365995, 0.0106411],
    "Flattop": [0.21557895, 0.41663158, 0.277263158, 0.083578947, 0.006947368],
}


def _cosine_series(n, N, coeffs):
    out = coeffs[0]
    for i in range(1, len(coeffs)):
        out = out + (-1) ** i * coeffs[i] * np.cos(i * 2 * np.pi * n / (N - 1))
    return out


def _win_1d(N, type_, param):
    """Symmetric 1-D window of length N; kwave/utils/signals.py:142-234."""
    n = np.arange(0, N)
    if type_ == "Bartlett":
        return 2 / (N - 1) * ((N - 1) / 2 - abs(n - (N - 1) / 2))
    if type_ == "Bartlett-Hanning":
        return 0.62 - 0.48 * abs(n / (N - 1) - 1 / 2) - 0.38 * np.cos(2 * np.pi * n / (N - 1))
    if type_ == "Blackman":
        return _cosine_series(n, N, [(1 - param) / 2, 0.5, param / 2])
    if type_ in _COSINE_SERIES:
        return _cosine_series(n, N, _COSINE_SERIES[type_])
    if type_ == "Cosine":
        return np.cos(np.pi * n / (N - 1) - np.pi / 2)
    if type_ == "Gaussian":
        return np.exp(-0.5 * ((n - (N - 1) / 2) / (param * (N - 1) / 2)) ** 2)
    if type_ == "Hamming":
        return 0.54 - 0.46 * np.cos(2 * np.pi * n / (N - 1))
    if type_ == "Hanning":
        return 0.5 - 0.5 * np.cos(2 * np.pi * n / (N - 1))
    if type_ == "Rectangular":
        return np.ones(N)
    if type_ == "Triangular":
        return 2 / N * (N / 2 - abs(n - (N - 1) / 2))
    if type_ == "Tukey":
        win = np.ones(N)
        idx = np.arange(0, (N - 1) * param / 2 + 1e-8)
        width = param * N
        win[: len(idx)] = 0.5 * (1 + np.cos(2 * np.pi / width * (idx - width / 2)))
        win[np.arange(-1, -len(idx) - 1, -1)] = win[: len(idx)]
        return win
    raise ValueError(f"Unknown window type: {type_}")


_DEFAULT_PARAM = {"Tukey": (0.5, 0, 1), "Blackman": (0.16, 0, 1), "Gaussian": (0.5, 0, 0.5), "Kaiser": (3, 0, 100)}


def get_win(N, type_, plot_win=False, param=None, rotation=False, symmetric=True, square=False):
    """Frequency-domain window of 1, 2 or 3 dimensions; returns (win, coherent_gain).

    Non-symmetric windows are built one sample longer and trimmed
    (signals.py:134, 246-248); 2-D/3-D windows come from rotating the 1-D
    window (``rotation=True``) or from outer products.
    """
    if plot_win:
        raise NotImplementedError("Plotting is not implemented.")
    if square:
        raise NotImplementedError("square windows are not used on the solver path")
    N = np.array(N, dtype=int)
    N = N if N.size > 1 else N.item()
    sym = np.array(symmetric, dtype=bool)
    if type_ in _DEFAULT_PARAM:
        d, lo, hi = _DEFAULT_PARAM[type_]
        param = np.clip(d if param is None else param, lo, hi)
    Nb = N + 1 * (1 - sym.astype(int))
    if np.size(Nb) == 1:
        Nb = int(Nb)
        win = _win_1d(Nb, type_, param)
        Nt = Nb if bool(sym) else Nb - 1
        win = win[:Nt][:, None]
        return win, win.sum() / Nt
    Nb = np.atleast_1d(Nb)
    if rotation:
        L = int(Nb.max())
        lin = np.squeeze(get_win(L, type_, param=param)[0])
        radius = (L - 1) / 2
        ll = np.linspace(-radius, radius, L)
        grids = np.meshgrid(*[np.linspace(-radius, radius, int(n)) for n in Nb], indexing="ij")
        r = np.sqrt(sum(g ** 2 for g in grids))
        r[r > radius] = radius
        win = interp1d(ll, lin)(r)
    else:
        parts = [np.squeeze(get_win(int(n), type_, param=param)[0]) for n in Nb]
        win = parts[0]
        for p in parts[1:]:
            win = np.multiply.outer(win, p)
    Nt = Nb - 1 * (1 - np.broadcast_to(sym, Nb.shape).astype(int))
    win = win[tuple(slice(0, int(n)) for n in Nt)]
    return win, win.sum() / np.prod(Nt)


def tone_burst(sample_freq, signal_freq, num_cycles, envelope="Gaussian", plot_signal=False, signal_length=0, signal_offset=0):
    """Enveloped tone burst, one row per offset; kwave/utils/signals.py:323-434."""
    from ..kgrid import _matlab_rem

    if plot_signal:
        raise NotImplementedError
    dt = 1 / sample_freq
    tone_length = num_cycles / signal_freq
    if _matlab_rem(tone_length, dt) < 1e-18:
        tone_t = np.linspace(0, tone_length, int(tone_length / dt) + 1)
    else:
        tone_t = np.arange(0, tone_length, dt)
    burst = np.sin(2 * np.pi * signal_freq * tone_t)
    tone_index = np.round(signal_offset)
    if isinstance(envelope, (list, np.ndarray)):
        up, down = envelope
        if num_cycles < up + down:
            raise ValueError("Input num_cycles must be longer than num_ring_up_cycles + num_ring_down_cycles.")
        period = 1 / signal_freq
        nu, nd = round(up * period / dt), round(down * period / dt)
        ua = np.arange(0, np.pi + 1e-8, np.pi / (nu - 1))
        da = np.arange(0, np.pi + 1e-8, np.pi / (nd - 1))
        burst[:nu] *= (-np.cos(ua) + 1) * 0.5
        burst[-nd:] *= (np.cos(da) + 1) * 0.5
    else:
        if envelope == "Gaussian":
            x_lim = 3
            wx = np.arange(-x_lim, x_lim + 1e-8, 2 * x_lim / (len(burst) - 1))
            burst = burst * gaussian(wx, 1, 0, 1)
            burst = burst * np.squeeze(get_win(len(burst), type_="Tukey", param=0.05)[0])
        elif envelope == "Rectangular":
            pass
        else:
            raise ValueError(f"Unknown envelope {envelope}.")
    offs = np.array(signal_offset)
    length = max(signal_length, int(offs.max()) + len(burst))
    sig = np.zeros((np.atleast_1d(offs).size, length))
    ti = np.atleast_1d(tone_index)
    if ti.size == 1:
        i0 = int(np.squeeze(ti))
        sig[:, i0:i0 + len(burst)] = burst
    else:
        for i, idx in enumerate(ti):
            sig[i, int(idx):int(idx) + len(burst)] = burst
    return sig
