"""The truncated 2-D DFT matrices of the surrogate's spectral layers against numpy's FFTs."""
import numpy as np

from kwave_amd.surrogate import dft_mats


def _apply(n, md, v):
    mx, my, _, _ = [a.astype(np.float64) for a in dft_mats(n, md)]
    bx = (mx @ v).reshape(2 * md, 2 * n)
    x = bx @ my
    return x[:, 0::2] + 1j * x[:, 1::2]


def _inverse(n, md, y):
    _, _, ny, nx = [a.astype(np.float64) for a in dft_mats(n, md)]
    yr = np.stack([y.real, y.imag], -1).reshape(2 * md, 2 * md)
    c = (yr @ ny).reshape(4 * md, n)
    return nx @ c


def test_forward_matches_rfft2_at_retained_modes():
    rng = np.random.default_rng(0)
    for n, md in ((64, 16), (32, 5)):
        v = rng.standard_normal((n, n))
        ref = np.fft.rfft2(v)
        kx = np.r_[0:md, n - md:n]
        want = ref[kx][:, :md]
        got = _apply(n, md, v)
        assert np.abs(got - want).max() < 1e-5 * np.abs(want).max()


def test_inverse_matches_irfft2_of_zero_padded_spectrum():
    rng = np.random.default_rng(1)
    for n, md in ((64, 16), (32, 5)):
        y = rng.standard_normal((2 * md, md)) + 1j * rng.standard_normal((2 * md, md))
        full = np.zeros((n, n // 2 + 1), complex)
        full[:md, :md], full[n - md:, :md] = y[:md], y[md:]
        want = np.fft.irfft2(full, s=(n, n))
        got = _inverse(n, md, y)
        assert np.abs(got - want).max() < 1e-5 * np.abs(want).max()


def test_round_trip_keeps_band_limited_field():
    # a field made of retained modes with ky > 0 is reproduced by forward then
    # inverse (the ky = 0 column of a real field pairs kx with -kx, and +md is
    # not retained)
    rng = np.random.default_rng(2)
    n, md = 64, 16
    y = rng.standard_normal((2 * md, md)) + 1j * rng.standard_normal((2 * md, md))
    y[:, 0] = 0
    full = np.zeros((n, n // 2 + 1), complex)
    full[:md, :md], full[n - md:, :md] = y[:md], y[md:]
    v = np.fft.irfft2(full, s=(n, n))
    back = _inverse(n, md, _apply(n, md, v))
    assert np.abs(back - v).max() < 1e-5 * np.abs(v).max()
