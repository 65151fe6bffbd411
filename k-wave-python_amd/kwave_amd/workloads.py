"""Seeded synthetic inputs for the benchmark and parity workloads.

Each builder returns ``(kgrid, medium, source, sensor, pml_size, pml_alpha)``
ready for ``kspaceFirstOrder2D/3D`` (PML inside the grid).  Every random
choice goes through ``numpy.random.default_rng(seed)``; the definitions are
frozen here so results are comparable across runs.

* ``ivp_2d``      128^2 homogeneous lossless initial-value problem, 50-point
                  Cartesian circle sensor (Nt = 664).
* ``hetero_2d``   512^2 heterogeneous absorbing medium (8 smoothed elliptical
                  inclusions), 64-cell additive line source (1 MHz, 3-cycle
                  Hann tone burst, 1 MPa), 256-cell line sensor, 2000 steps.
* ``nonlinear_3d`` 512^3 Gaussian-random c/rho, B/A = 6, power-law absorption,
                  16x16 planar source (0.5 MHz, 4 cycles, 2 MPa), 256x256
                  planar sensor recording p and p_max, 1500 steps.
* ``scan_line_2d`` / ``scan_sweep_2d``  uneven 2-D linear-array sweep
                  (96 instances of 256^2 .. 640^2, 1200-2600 steps).
* ``surrogate_2d`` 128^2 sound-speed map -> 32 sensor traces of 600 steps: the
                  samples of the learned surrogate and the adjoint problem.
The grid edge ``n`` and step count ``nt`` arguments shrink the problems for
parity tests.
"""
from __future__ import annotations

import numpy as np

from .data import Vector
from .kgrid import kWaveGrid
from .kmedium import kWaveMedium
from .ksensor import kSensor
from .ksource import kSource
from .utils.mapgen import make_cart_circle, make_disc
from .utils.signals import get_win, tone_burst


def _gaussian_filter_periodic(a, sigma):
    """Periodic convolution with a unit-area Gaussian of std ``sigma`` cells (spectral)."""
    import scipy.fft as sfft

    spec = sfft.rfftn(a, workers=-1)
    shape = a.shape
    k2 = 0
    for ax, n in enumerate(shape):
        f = sfft.rfftfreq(n) if ax == len(shape) - 1 else sfft.fftfreq(n)
        sh = [1] * len(shape)
        sh[ax] = f.size
        k2 = k2 + ((2 * np.pi * f) ** 2).reshape(sh)
    spec *= np.exp(-0.5 * sigma ** 2 * k2)
    return sfft.irfftn(spec, s=shape, workers=-1)


def _affine(a, lo, hi):
    mn, mx = a.min(), a.max()
    return lo + (a - mn) * (hi - lo) / (mx - mn)


def ivp_2d(seed=1, n=128, pml=20):
    """128^2 homogeneous IVP: two smoothed discs, Cartesian circle sensor."""
    rng = np.random.default_rng(seed)
    dx = 1e-4
    kgrid = kWaveGrid(Vector([n, n]), Vector([dx, dx]))
    c = 1500.0
    dt = 0.3 * dx / c
    t_end = 1.1 * np.sqrt(2) * n * dx / c
    kgrid.setTime(int(np.floor(t_end / dt) + 1), dt)
    medium = kWaveMedium(sound_speed=c, density=1000.0)
    p0 = np.zeros((n, n))
    for r, amp in ((5, 5.0), (3, 3.0)):
        cx, cy = rng.integers(pml + r + 3, n - pml - r - 2, size=2)
        p0 = p0 + amp * make_disc(Vector([n, n]), Vector([int(cx), int(cy)]), r)
    source = kSource()
    source.p0 = p0
    sensor = kSensor(make_cart_circle(4e-3 * n / 128, 50))
    sensor.record = ["p"]
    return kgrid, medium, source, sensor, (pml, pml), (2.0, 2.0)


def hetero_2d(seed=2, n=512, nt=2000, pml=20):
    """512^2 heterogeneous absorbing medium with a time-varying line source."""
    rng = np.random.default_rng(seed)
    s = n / 512
    dx = 5e-5
    c = np.full((n, n), 1500.0)
    rho = np.full((n, n), 1000.0)
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    lo, hi = pml + int(40 * s), n - pml - int(40 * s)
    for _ in range(8):
        cx, cy = rng.uniform(lo, hi, size=2)
        a, b = rng.uniform(8 * s, 40 * s, size=2)
        th = rng.uniform(0, np.pi)
        cv, rv = rng.uniform(1400, 1800), rng.uniform(1000, 1200)
        x, y = ii - cx, jj - cy
        u = x * np.cos(th) + y * np.sin(th)
        v = -x * np.sin(th) + y * np.cos(th)
        inside = (u / a) ** 2 + (v / b) ** 2 <= 1
        c[inside] = cv
        rho[inside] = rv
    from scipy.ndimage import gaussian_filter

    c = gaussian_filter(c, 2 * max(s, 0.25), mode="reflect")
    rho = gaussian_filter(rho, 2 * max(s, 0.25), mode="reflect")
    kgrid = kWaveGrid(Vector([n, n]), Vector([dx, dx]))
    dt = 0.3 * dx / c.max()
    kgrid.setTime(int(nt), dt)
    medium = kWaveMedium(sound_speed=c, density=rho, alpha_coeff=0.75, alpha_power=1.5)
    f0 = 1e6
    nsig = int(round(3 / (f0 * dt)))
    t = np.arange(nsig) * dt
    sig = 1e6 * np.sin(2 * np.pi * f0 * t) * np.squeeze(get_win(nsig, "Hanning")[0])
    source = kSource()
    mask = np.zeros((n, n), dtype=bool)
    x0 = int(round(30 * s))
    mask[x0, n // 2 - int(32 * s): n // 2 + int(32 * s)] = True
    source.p_mask = mask
    source.p = sig[None, :]
    smask = np.zeros((n, n), dtype=bool)
    smask[n - 1 - int(round(30 * s)), n // 4: 3 * n // 4] = True
    sensor = kSensor(smask)
    sensor.record = ["p"]
    return kgrid, medium, source, sensor, (pml, pml), (2.0, 2.0)


def nonlinear_3d(seed=3, n=512, nt=1500, pml=16):
    """512^3 heterogeneous nonlinear absorbing medium, planar source and sensor."""
    rng = np.random.default_rng(seed)
    s = n / 512
    dx = 1e-4
    shape = (n, n, n)
    sig_cells = 8 * s
    g1 = _gaussian_filter_periodic(rng.standard_normal(shape, dtype=np.float32), sig_cells)
    g2 = _gaussian_filter_periodic(rng.standard_normal(shape, dtype=np.float32), sig_cells)
    c = _affine(g1.astype(np.float64), 1450.0, 1650.0)
    rho = _affine(g2.astype(np.float64), 950.0, 1100.0)
    del g1, g2
    kgrid = kWaveGrid(Vector([n, n, n]), Vector([dx, dx, dx]))
    dt = 0.3 * dx / 1650.0
    kgrid.setTime(int(nt), dt)
    medium = kWaveMedium(sound_speed=c, density=rho, alpha_coeff=0.5, alpha_power=1.1, BonA=6.0)
    source = kSource()
    mask = np.zeros(shape, dtype=bool)
    h = max(int(8 * s), 1)
    mask[int(round(24 * s)), n // 2 - h: n // 2 + h, n // 2 - h: n // 2 + h] = True
    source.p_mask = mask
    source.p = 2e6 * tone_burst(1 / dt, 0.5e6, 4)
    smask = np.zeros(shape, dtype=bool)
    smask[n // 2, n // 4: 3 * n // 4, n // 4: 3 * n // 4] = True
    sensor = kSensor(smask)
    sensor.record = ["p", "p_max"]
    return kgrid, medium, source, sensor, (pml, pml, pml), (2.0, 2.0, 2.0)


def nonlinear_3d_box(seed, shape, nt=300, long_pml=16, short_pml=4):
    """The 512^3 workload's physics on an elongated box (parity cases of the 512-point kernels).

    Gaussian-random c in [1450, 1650] and rho in [950, 1100] (correlation 8
    cells), B/A = 6, power-law 0.5 dB/(MHz^1.1 cm); dx = 0.1 mm, dt = 0.3 dx /
    1650.  The longest axis carries a 16-cell PML, the others 4.  An additive
    8 x 8 plane source 24 cells along the long axis emits a 1.5 MHz, 3-cycle
    tone burst of 2 MPa (shorter than nt, so the steps after it has ended run
    too); a full cross-section sensor plane 60 cells further (or just inside the
    far PML of a shorter box) records p and p_max.
    """
    rng = np.random.default_rng(seed)
    shape = tuple(int(n) for n in shape)
    ax = int(np.argmax(shape))
    dx = 1e-4
    c = _affine(_gaussian_filter_periodic(rng.standard_normal(shape), 8.0), 1450.0, 1650.0)
    rho = _affine(_gaussian_filter_periodic(rng.standard_normal(shape), 8.0), 950.0, 1100.0)
    kgrid = kWaveGrid(Vector(list(shape)), Vector([dx] * 3))
    dt = 0.3 * dx / 1650.0
    kgrid.setTime(int(nt), dt)
    medium = kWaveMedium(sound_speed=c, density=rho, alpha_coeff=0.5, alpha_power=1.1, BonA=6.0)
    src_sl = [slice(n // 2 - 4, n // 2 + 4) for n in shape]
    src_sl[ax] = 24
    mask = np.zeros(shape, dtype=bool)
    mask[tuple(src_sl)] = True
    source = kSource()
    source.p_mask = mask
    source.p = 2e6 * tone_burst(1 / dt, 1.5e6, 3)
    sen_sl = [slice(None)] * 3
    sen_sl[ax] = min(84, shape[ax] - long_pml - 2)
    smask = np.zeros(shape, dtype=bool)
    smask[tuple(sen_sl)] = True
    sensor = kSensor(smask)
    sensor.record = ["p", "p_max"]
    pml = tuple(long_pml if i == ax else short_pml for i in range(3))
    return kgrid, medium, source, sensor, pml, (2.0, 2.0, 2.0)


SCAN_SIZES = ((256, 256), (384, 512), (512, 768), (640, 640))


def scan_line_2d(seed, shape=None, nt=None, pml=20):
    """One linear-array scan-line simulation of the uneven 2-D sweep.

    Speckle phantom: c = 1540 (1 + U[-0.01, 0.01]) white noise plus 0-4
    hyperechoic discs (radius U[10, 40] cells, +5 % c), rho = 1000, lossless.
    Transmit: 32 elements x 2 cells along y at x = pml + 10, 2 MHz 3-cycle tone
    burst per element, delayed to focus 30 mm deep on the array axis (one signal
    row per mask point, C order).  Receive: the 32 element centres.  dx = 50 um,
    Nt uniform in [1200, 2600], PML 20.
    """
    rng = np.random.default_rng(seed)
    if shape is None:
        shape = SCAN_SIZES[rng.integers(len(SCAN_SIZES))]
    nx, ny = shape
    dx = 5e-5
    c = 1540.0 * (1.0 + rng.uniform(-0.01, 0.01, size=shape))
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    for _ in range(int(rng.integers(0, 5))):
        r = rng.uniform(10, 40) * min(1.0, min(nx, ny) / 256)
        cx, cy = rng.uniform(pml + r, nx - pml - r), rng.uniform(pml + r, ny - pml - r)
        c[(ii - cx) ** 2 + (jj - cy) ** 2 <= r * r] *= 1.05
    if nt is None:
        nt = int(rng.integers(1200, 2601))
    kgrid = kWaveGrid(Vector([nx, ny]), Vector([dx, dx]))
    dt = 0.3 * dx / c.max()
    kgrid.setTime(int(nt), dt)
    medium = kWaveMedium(sound_speed=c, density=1000.0)
    x0 = pml + 10
    y_start = ny // 2 - 32
    mask = np.zeros(shape, dtype=bool)
    mask[x0, y_start:y_start + 64] = True
    el_y = (y_start + 2 * np.arange(32) + 0.5) * dx
    focus = np.array([x0 * dx + 30e-3, (ny // 2) * dx])
    dist = np.hypot(focus[0] - x0 * dx, focus[1] - el_y)
    delay = np.round((dist.max() - dist) / 1540.0 / dt).astype(int)
    burst = tone_burst(1 / dt, 2e6, 3, signal_offset=delay)  # one row per element
    source = kSource()
    source.p_mask = mask
    source.p = 1e5 * np.repeat(burst, 2, axis=0)  # two cells per element, C order along y
    smask = np.zeros(shape, dtype=bool)
    smask[x0, y_start + 2 * np.arange(32)] = True
    sensor = kSensor(smask)
    sensor.record = ["p"]
    return kgrid, medium, source, sensor, (pml, pml), (2.0, 2.0)


def scan_sweep_2d(seed=4, count=96):
    """The 96-instance uneven sweep: instance i uses seed 1000 * seed + i."""
    return [scan_line_2d(1000 * seed + i) for i in range(count)]


SURROGATE_C_RANGE = (1400.0, 1650.0)


def surrogate_c_map(seed, n=128):
    """Sound-speed map of one surrogate sample.

    Gaussian random field (white noise, periodic Gaussian filter with a
    correlation length of 10 cells) mapped affinely onto [1400, 1650] m/s, plus
    0-3 ellipses (semi-axes 6-20 cells, random angle) at a uniform value in the
    same range, then a 1-cell Gaussian smoothing.
    """
    rng = np.random.default_rng(seed)
    lo, hi = SURROGATE_C_RANGE
    c = _affine(_gaussian_filter_periodic(rng.standard_normal((n, n)), 10.0 * n / 128), lo, hi)
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    s = n / 128
    for _ in range(int(rng.integers(0, 4))):
        cx, cy = rng.uniform(30 * s, n - 20 * s, size=2)
        a, b = rng.uniform(6 * s, 20 * s, size=2)
        th = rng.uniform(0, np.pi)
        x, y = ii - cx, jj - cy
        u = x * np.cos(th) + y * np.sin(th)
        v = -x * np.sin(th) + y * np.cos(th)
        c[(u / a) ** 2 + (v / b) ** 2 <= 1] = rng.uniform(lo, hi)
    from scipy.ndimage import gaussian_filter

    return np.clip(gaussian_filter(c, 1.0, mode="reflect"), lo, hi)


def surrogate_2d(seed, n=128, nt=600, c=None):
    """One surrogate sample: c map (seeded, or given) and the fixed acquisition.

    128^2, dx = 0.1 mm, rho = 1000, lossless.  dt = 0.3 dx / 1650 (the top of the
    c range, so dt and Nt do not depend on the sample).  Source: 32-cell line at
    x = 24, y in [48, 80), additive, tone_burst(1/dt, 1.5 MHz, 3 cycles) x 1e5.
    Sensor: the same 32 cells, recording p.  PML 20, alpha 2.
    """
    s = n / 128
    dx = 1e-4
    if c is None:
        c = surrogate_c_map(seed, n)
    kgrid = kWaveGrid(Vector([n, n]), Vector([dx, dx]))
    dt = 0.3 * dx / SURROGATE_C_RANGE[1]
    kgrid.setTime(int(nt), dt)
    medium = kWaveMedium(sound_speed=np.asarray(c, dtype=np.float64), density=1000.0)
    mask = np.zeros((n, n), dtype=bool)
    mask[int(round(24 * s)), int(round(48 * s)):int(round(80 * s))] = True
    source = kSource()
    source.p_mask = mask
    source.p = 1e5 * tone_burst(1 / dt, 1.5e6, 3)
    sensor = kSensor(mask.copy())
    sensor.record = ["p"]
    pml = max(2, int(round(20 * s)))
    return kgrid, medium, source, sensor, (pml, pml), (2.0, 2.0)
