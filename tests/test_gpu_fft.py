"""Transform kernels vs numpy.fft: real(ifftn(op * fftn(f))) for every operator kind.

This pins the HIP FFT passes (Stockham radices 2/3/4/5/7/8, the paired
real<->half-complex contiguous-axis transform and the strided column passes)
against numpy.fft (pocketfft), the library the reference solver calls at
kwave/solvers/kspace_solver.py:702,705,796.  Tolerance: relative L2 < 2e-6
(FP32 transforms of O(1) data).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ref(f, spacing, kind, axis, kap, use_sg, power):
    nd = f.ndim
    ks = []
    for a in range(nd):
        sh = [1] * nd
        sh[a] = f.shape[a]
        ks.append((2 * np.pi * np.fft.fftfreq(f.shape[a], d=spacing[a])).reshape(sh))
    kmag = np.sqrt(sum(k ** 2 for k in ks))
    kappa = np.sinc(kap * kmag / np.pi) if kap > 0 else 1.0
    k, d = ks[axis], spacing[axis]
    if kind == 0:
        op = np.ones_like(kmag)
    elif kind == 1:
        op = 1j * k * (np.exp(1j * k * d / 2) if use_sg else 1) * kappa
    elif kind == 2:
        op = 1j * k * (np.exp(-1j * k * d / 2) if use_sg else 1) * kappa
    elif kind == 3:
        with np.errstate(divide="ignore"):
            op = np.where(kmag == 0, 0, kmag ** power)
    elif kind == 5:
        op = np.cos(kap * kmag) if kap > 0 else np.ones_like(kmag)
    elif kind == 6:
        op = np.exp(-1j * k * d / 2)
    return np.real(np.fft.ifftn(op * np.fft.fftn(f)))


SHAPES_2D = [(64, 64), (128, 96), (48, 160), (384, 512), (640, 56), (24, 448)]
SHAPES_3D = [(32, 48, 64), (16, 64, 40), (64, 64, 64)]


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.parametrize("shape", SHAPES_2D + SHAPES_3D)
@pytest.mark.parametrize("kind,axis", [(0, 0), (1, 0), (1, -1), (2, 0), (2, -1), (3, 0), (5, 0), (6, 0), (6, -1)])
def test_spectral_apply(kwlib, shape, kind, axis):
    from kwave_amd import _abi

    rng = np.random.default_rng(hash((shape, kind, axis)) % 2**32)
    f = rng.standard_normal(shape)
    spacing = [1e-4 * (1 + 0.1 * i) for i in range(len(shape))]
    ax = axis % len(shape)
    # kap = c_ref dt / 2 for a CFL-0.3 step, so kap * max|k| is about 0.5
    kap = 0.15 * spacing[0] if kind in (1, 2, 5) else 0.0
    power = -0.5 if kind == 3 else 0.0
    got = _abi.spectral_apply(f, spacing, kind, axis=ax, kap=kap, power=power)
    want = _ref(f, spacing, kind, ax, kap, True, power)
    assert _rel(got, want) < 2e-6, (shape, kind, ax, _rel(got, want))
