"""Plane-wave physics checks shared by the oracle pins and the GPU tests.

Each case builds a 2-D problem whose field is uniform along y (a plane wave
along x, PML only along x), runs it through a ``runner`` (the FP64 oracle or
the HIP solver) and measures a quantity that the physics of the k-Wave model
fixes independently of any implementation:

* ``attenuation``: a broadband pulse recorded at two planes; the amplitude
  ratio per frequency must follow exp(-alpha_0 w^y dx) (alpha_0 from
  db2neper) and the phase speed the Kramers-Kronig relation for power-law
  media (kwave/utils/mapgen.py:328-358 with w0 -> 0 for y > 1):
  1/c(w) = 1/c0 + alpha_0 tan(pi y / 2) w^(y-1).
* ``nonlinearity``: a continuous wave driven by a Dirichlet line; the second
  harmonic at two planes, inverted through the Fubini solution
  B_n = 2 J_n(n s)/(n s), gives the growth of the shock parameter s, hence
  beta = (s2 - s1) rho0 c0^3 / (w P0 dx), which must equal 1 + B/(2A).
* ``intensity``: a steady plane wave; the time-averaged intensity the solver
  derives from its p and u records must equal P^2 / (2 rho0 c0).
* ``near_unity``: the issue-664 case of the reference's tests
  (tests/test_issue_664_alpha_power_near_unity.py:58-83): heterogeneous
  absorption with alpha_power close to 1 must stay finite with
  alpha_mode='no_dispersion'.

``runner(kgrid, medium, source, sensor, pml_size, pml_alpha)`` returns the
result dict of ``Simulation.run()``.
"""
from __future__ import annotations

import numpy as np

C0, RHO0 = 1500.0, 1000.0


def _grid(nx, ny, dx, dt, nt):
    from kwave_amd.kgrid import kWaveGrid

    kg = kWaveGrid([nx, ny], [dx, dx])
    kg.setTime(int(nt), float(dt))
    return kg


def _line_source(nx, ny, x):
    m = np.zeros((nx, ny), dtype=bool)
    m[x, :] = True
    return m


def _points(nx, ny, xs):
    m = np.zeros((nx, ny), dtype=bool)
    for x in xs:
        m[x, 0] = True
    return m


def db2neper(alpha, y):
    """dB/(MHz^y cm) -> Np/((rad/s)^y m) (kwave/utils/conversion.py:34)."""
    return 100.0 * alpha * (1e-6 / (2 * np.pi)) ** y / (20 * np.log10(np.e))


# ---------------------------------------------------------------------------
def attenuation(runner, alpha_db, y, *, ny=4, cfl=0.3, band=(0.8e6, 2.3e6)):
    """Measured vs predicted attenuation and dispersion across a frequency band.

    Returns (amp_err, disp_err): amp_err = measured / predicted alpha(w) - 1 per
    frequency; disp_err = (measured - predicted) (Re k - w/c0), divided by the
    predicted alpha(w) (the dispersion term is of the size of alpha).
    """
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource

    nx, dx = 512, 5e-5
    dt = cfl * dx / C0
    xs, x1, x2 = 30, 80, 430
    nt = int((x2 - xs + 60) * dx / C0 / dt + 4e-6 / dt)
    kg = _grid(nx, ny, dx, dt, nt)
    med = kWaveMedium(sound_speed=C0, density=RHO0, alpha_coeff=alpha_db, alpha_power=y)
    src = kSource()
    src.p_mask = _line_source(nx, ny, xs)
    t = np.arange(nt) * dt
    f0, t0, sd = 1.5e6, 1.5e-6, 0.35e-6
    src.p = (np.sin(2 * np.pi * f0 * (t - t0)) * np.exp(-0.5 * ((t - t0) / sd) ** 2))[None, :]
    sen = kSensor(_points(nx, ny, (x1, x2)))
    sen.record = ["p"]
    p = runner(kg, med, src, sen, (20, 0), (2.0, 0.0))["p"]
    P1, P2 = np.fft.rfft(p[0]), np.fft.rfft(p[1])
    w = 2 * np.pi * np.fft.rfftfreq(nt, dt)
    sel = (w > 2 * np.pi * band[0]) & (w < 2 * np.pi * band[1])
    w, P1, P2 = w[sel], P1[sel], P2[sel]
    D = (x2 - x1) * dx
    a0 = db2neper(alpha_db, y)
    alpha_th = a0 * w ** y
    alpha_meas = -np.log(np.abs(P2) / np.abs(P1)) / D
    # phase of P2/P1 relative to lossless propagation: -(Re k - w/c0) D
    dk_meas = -np.angle(P2 / P1 * np.exp(1j * w * D / C0)) / D
    dk_th = a0 * np.tan(np.pi * y / 2) * w ** y if abs(y - 2.0) > 1e-9 else np.zeros_like(w)
    return alpha_meas / alpha_th - 1.0, (dk_meas - dk_th) / alpha_th


def kk_phase_speed(w, c0, a0, y):
    """Kramers-Kronig phase speed about w0 -> 0 (kwave/utils/mapgen.py:328-358, y > 1)."""
    return 1.0 / (1.0 / c0 + a0 * np.tan(y * np.pi / 2) * w ** (y - 1))


# ---------------------------------------------------------------------------
def _fubini_ratio(s):
    """|B_2 / B_1| of the Fubini solution at shock parameter s < 1."""
    from scipy.special import jv

    return (jv(2, 2 * s) / s) / (2 * jv(1, s) / s)


def nonlinearity(runner, bona, *, ny=4, amp=1e6, f0=1e6):
    """Measured beta of a continuous plane wave between two planes (expect 1 + B/2A)."""
    from scipy.optimize import brentq
    from scipy.special import jv

    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource

    nx, dx, cfl = 768, 5e-5, 0.2
    dt = cfl * dx / C0
    per = 1.0 / f0
    xs, x1, x2, nramp = 30, 150, 650, 3
    nt = int(((x2 - xs) * dx / C0 + (nramp + 8) * per) / dt)
    kg = _grid(nx, ny, dx, dt, nt)
    med = kWaveMedium(sound_speed=C0, density=RHO0, BonA=bona)
    src = kSource()
    src.p_mask = _line_source(nx, ny, xs)
    src.p_mode = "dirichlet"
    t = np.arange(nt) * dt
    ramp = 0.5 - 0.5 * np.cos(np.pi * np.clip(t / (nramp * per), 0, 1))
    src.p = (amp * ramp * np.sin(2 * np.pi * f0 * t))[None, :]
    sen = kSensor(_points(nx, ny, (x1, x2)))
    sen.record = ["p"]
    p = runner(kg, med, src, sen, (20, 0), (2.0, 0.0))["p"]
    n4 = int(round(4 * per / dt))  # four whole periods of the steady state
    s = []
    for i, xi in enumerate((x1, x2)):
        i0 = int(((xi - xs) * dx / C0 + (nramp + 1) * per) / dt)
        seg, tt = p[i, i0:i0 + n4], t[i0:i0 + n4]
        h1 = abs(2 * np.mean(seg * np.exp(-2j * np.pi * f0 * tt)))
        h2 = abs(2 * np.mean(seg * np.exp(-4j * np.pi * f0 * tt)))
        s.append((brentq(lambda v: _fubini_ratio(v) - h2 / h1, 1e-6, 0.99), h1))
    (s1, h1), (s2, _) = s
    p0 = h1 / (2 * jv(1, s1) / s1)
    return (s2 - s1) * RHO0 * C0 ** 3 / (2 * np.pi * f0 * p0 * (x2 - x1) * dx)


# ---------------------------------------------------------------------------
def intensity(runner, *, ny=4, amp=1e5, f0=1e6, spp=30, ppw=30):
    """Ix_avg and Iy_avg of a steady plane wave along x, in units of P^2 / (2 rho0 c0) (expect 1 and 0)."""
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource

    nx = 256
    dx, dt, per = C0 / f0 / ppw, 1.0 / f0 / spp, 1.0 / f0
    xs, x1, nramp = 30, 120, 2
    start = int(((x1 - xs) * dx / C0 + (nramp + 2) * per) / dt)
    nt = start + 6 * spp  # six whole periods recorded
    kg = _grid(nx, ny, dx, dt, nt)
    med = kWaveMedium(sound_speed=C0, density=RHO0)
    src = kSource()
    src.p_mask = _line_source(nx, ny, xs)
    src.p_mode = "dirichlet"
    t = np.arange(nt) * dt
    ramp = 0.5 - 0.5 * np.cos(np.pi * np.clip(t / (nramp * per), 0, 1))
    src.p = (amp * ramp * np.sin(2 * np.pi * f0 * t))[None, :]
    sen = kSensor(_points(nx, ny, (x1,)))
    sen.record = ["p", "I_avg"]
    sen.record_start_index = start + 1
    r = runner(kg, med, src, sen, (20, 0), (2.0, 0.0))
    P2 = 2 * np.mean(np.asarray(r["p"][0], dtype=np.float64) ** 2)
    i_ref = P2 / (2 * RHO0 * C0)
    return float(r["Ix_avg"][0]) / i_ref, float(r["Iy_avg"][0]) / i_ref


# ---------------------------------------------------------------------------
def near_unity_case(alpha_power, alpha_mode="no_dispersion"):
    """The 64 x 64 heterogeneous issue-664 problem (reference tests/test_issue_664_alpha_power_near_unity.py:58-83).

    Returns (kgrid, medium, source, sensor) for kspaceFirstOrder(pml_inside=False, pml_alpha=10, smooth_p0=False).
    """
    from scipy.signal import gausspulse

    from kwave_amd.kgrid import kWaveGrid
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource

    N = (64, 64)
    c = 1420 + 220 * np.random.default_rng(0).random(N)
    atten = 0.0025 + 6e-2 * np.abs(c - c[0, 0])
    kg = kWaveGrid(list(N), [0.15e-3, 0.15e-3])
    kg.makeTime(c)
    med = kWaveMedium(sound_speed=c, density=1000 * np.ones(N), alpha_coeff=atten, alpha_power=alpha_power,
                      alpha_mode=alpha_mode)
    src = kSource()
    src.p_mask = np.zeros(N, dtype=bool)
    src.p_mask[N[0] // 4, N[1] // 4] = True
    t = kg.t_array.squeeze()
    src.p = gausspulse(t - 5e-7, fc=1e6, bw=0.75)[np.newaxis, :]
    src.p_mode = "dirichlet"
    return kg, med, src, kSensor(mask=np.ones(N, dtype=bool))
