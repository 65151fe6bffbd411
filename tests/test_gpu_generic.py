"""Axis lengths without fused kernels, on the generic-length family (csrc/kw_generic.hip).

The reference transforms any length (numpy.fft, kwave/solvers/kspace_solver.py:702-705,
796), and the default PML-outside expansion of kspaceFirstOrder turns common grids
into such lengths (64 -> 104 = 8 * 13, 128 -> 168 = 8 * 3 * 7).  Checks:

* the transform kernels against numpy.fft at lengths with factors 2, 3, 5, 7, 13, 37,
  primes (97, 101) and the longest (4096), every operator kind, rel-L2 < 2e-6;
* full solver runs against the FP64 oracle (rel-L2 < 1e-4) on generic lengths, on odd
  row counts (the fused contiguous-axis kernels pair rows, so such grids also take the
  generic family), with every source mode, velocity records and Cartesian sensors;
* the reference's default kspaceFirstOrder call on 64^2 and 128^2, and the
  issue-664 near-unity alpha_power case (reference tests/test_issue_664_alpha_power_near_unity.py);
* a batch of generic-length instances equals the instances run one at a time.
"""
import copy

import numpy as np
import pytest

from test_gpu_fft import _ref, _rel
from test_gpu_parity import _both, _compare, _grid, _hetero, _p0, _run_oracle

pytestmark = pytest.mark.gpu

SHAPES = [(100, 120), (144, 200), (243, 336), (1000, 24), (24, 1000), (2048, 16), (16, 2048), (101, 128), (97, 35),
          (104, 104), (162, 296), (4096, 16), (16, 4096), (36, 44, 50), (27, 25, 33), (20, 104, 64)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kind,axis", [(0, 0), (1, 0), (1, -1), (2, 0), (3, 0), (5, 0), (6, -1)])
def test_generic_spectral_apply(shape, kind, axis):
    from kwave_amd import _abi

    for n in shape:
        assert _abi.size_supported(n)
    rng = np.random.default_rng(hash((shape, kind, axis)) % 2**32)
    f = rng.standard_normal(shape)
    spacing = [1e-4 * (1 + 0.1 * i) for i in range(len(shape))]
    ax = axis % len(shape)
    kap = 0.15 * spacing[0] if kind in (1, 2, 5) else 0.0
    power = -0.5 if kind == 3 else 0.0
    got = _abi.spectral_apply(f, spacing, kind, axis=ax, kap=kap, power=power)
    want = _ref(f, spacing, kind, ax, kap, True, power)
    assert _rel(got, want) < 2e-6, (shape, kind, ax, _rel(got, want))


# steps per case: the wave must still be inside the grid at the end, or p_final is a
# residual of ~1e-4 of the recorded peak and FP32 rounding alone reaches the 1e-4 bar
@pytest.mark.parametrize("shape,physics,nt", [((100, 120), "nonlinear_powerlaw", 90), ((101, 128), "stokes", 90),
                                              ((97, 35), "lossless", 90), ((36, 44, 50), "nonlinear_powerlaw", 90),
                                              ((27, 25, 33), "powerlaw", 30), ((27, 25, 33), "nonlinear_powerlaw", 30)])
def test_generic_physics_ivp(shape, physics, nt):
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource

    kg = _grid(shape, nt=nt)
    c, r = _hetero(shape, 3)
    kw = {"lossless": {}, "stokes": dict(alpha_coeff=0.75, alpha_power=2.0),
          "powerlaw": dict(alpha_coeff=0.75, alpha_power=1.5),
          "nonlinear_powerlaw": dict(alpha_coeff=0.5, alpha_power=1.1, BonA=6.0)}[physics]
    med = kWaveMedium(sound_speed=c, density=r, **kw)
    src = kSource()
    src.p0 = 1e5 * _p0(shape, 4)
    mask = np.zeros(shape, dtype=bool)
    mask[tuple(slice(s // 4, 3 * s // 4, 3) for s in shape)] = True
    sen = kSensor(mask)
    sen.record = ["p", "p_final", "p_max", "u", "u_final"]
    pml = tuple(min(8, s // 4) for s in shape)
    g, o = _both((kg, med, src, sen, pml, tuple([2.0] * len(shape))))
    _compare(g, o)


@pytest.mark.parametrize("shape", [(101, 72), (45, 104), (30, 33, 40)])
@pytest.mark.parametrize("mode,rows", [("additive", "single"), ("additive", "many"), ("dirichlet", "many"),
                                       ("additive-no-correction", "single")])
def test_generic_sources(shape, mode, rows):
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource

    kg = _grid(shape, nt=110)
    c, r = _hetero(shape, 5)
    med = kWaveMedium(sound_speed=c, density=r, alpha_coeff=0.6, alpha_power=1.3)
    src = kSource()
    mask = np.zeros(shape, dtype=bool)
    mid = tuple(s // 2 for s in shape)
    sl = list(mid)
    sl[0] = 10
    sl[1] = slice(mid[1] - 6, mid[1] + 6)
    mask[tuple(sl)] = True
    src.p_mask = mask
    npt = int(mask.sum())
    t = np.arange(70) * float(kg.dt)
    base = np.sin(2 * np.pi * 1e6 * t) * np.hanning(70)
    src.p = 1e3 * (base[None, :] if rows == "single" else np.stack([np.roll(base, k) for k in range(npt)]))
    src.p_mode = mode
    umask = np.zeros(shape, dtype=bool)
    usl = list(mid)
    usl[0] = shape[0] - 12
    umask[tuple(usl)] = True
    src.u_mask = umask
    src.ux = 1e-3 * base[None, :]
    src.u_mode = mode
    sen = kSensor(np.ones(shape, dtype=bool))
    sen.record = ["p", "ux_staggered", "p_final"]
    pml = tuple(min(8, s // 4) for s in shape)
    g, o = _both((kg, med, src, sen, pml, tuple([2.0] * len(shape))))
    _compare(g, o)


def test_generic_cartesian_sensor_and_intensity():
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource
    from kwave_amd.utils.mapgen import make_cart_circle

    shape = (90, 110)
    kg = _grid(shape, nt=140)
    src = kSource()
    src.p0 = _p0(shape, 71)
    sen = kSensor(make_cart_circle(3e-3, 40))
    sen.record = ["p", "u", "I_avg"]
    g, o = _both((kg, kWaveMedium(sound_speed=1500.0, density=1000.0), src, sen, (10, 10), (2.0, 2.0)))
    _compare(g, o)


@pytest.mark.parametrize("n", [64, 128])
def test_reference_default_unified_call(n):
    """kspaceFirstOrder with its defaults (pml_size 20 outside the grid): 64 -> 104, 128 -> 168 points."""
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.kspaceFirstOrder import _expand_for_pml_outside, kspaceFirstOrder
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource
    from kwave_amd.utils.filters import smooth

    kg = _grid((n, n), nt=100)
    src = kSource()
    src.p0 = _p0((n, n), 81)
    sen = kSensor(np.ones((n, n), dtype=bool))
    sen.record = ["p", "p_final"]
    med = kWaveMedium(sound_speed=1500.0, density=1000.0)
    g = kspaceFirstOrder(kg, med, copy.deepcopy(src), sen)
    kg2, med2, src2, sen2 = _expand_for_pml_outside(kg, med, copy.deepcopy(src), sen, (20, 20))
    assert tuple(kg2.N) == (n + 40, n + 40)
    src2.p0 = smooth(np.asarray(src2.p0, dtype=float), restore_max=True)
    o = _run_oracle(kg2, med2, src2, sen2, (20, 20), (2.0, 2.0), smooth_p0=False)
    _compare(g, o)


@pytest.mark.parametrize("alpha_power", [0.97, 0.99, 1.005, 1.03])
def test_near_unity_alpha_power_on_gpu(alpha_power):
    """The reference's issue-664 case (tests/test_issue_664_alpha_power_near_unity.py:86-107) through
    kspaceFirstOrder on the GPU: finite and bounded, and equal to the oracle."""
    import physics_cases as pc
    from kwave_amd.kspaceFirstOrder import _expand_for_pml_outside, kspaceFirstOrder

    kg, med, src, sen = pc.near_unity_case(alpha_power)
    got = kspaceFirstOrder(kg, med, copy.deepcopy(src), sen, pml_inside=False, smooth_p0=False, pml_alpha=10)
    p = np.asarray(got["p"])
    assert not np.any(np.isnan(p))
    assert np.nanmax(np.abs(p)) < 1e3
    kg2, med2, src2, sen2 = _expand_for_pml_outside(kg, med, copy.deepcopy(src), sen, (20, 20))
    want = _run_oracle(kg2, med2, src2, sen2, (20, 20), (10.0, 10.0), smooth_p0=False)
    _compare(got, want, keys=["p"])


def test_generic_batch_equals_looped():
    from kwave_amd.batch import BatchSimulation
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource
    from kwave_amd.solvers.kspace_solver import Simulation

    cases = []
    for i, (shape, nt) in enumerate([((100, 120), 80), ((104, 104), 95), ((100, 120), 70), ((101, 72), 60)]):
        kg = _grid(shape, nt=nt)
        c, _ = _hetero(shape, 90 + i)
        src = kSource()
        m = np.zeros(shape, dtype=bool)
        m[12, shape[1] // 2 - 4: shape[1] // 2 + 4] = True
        src.p_mask = m
        t = np.arange(40) * float(kg.dt)
        src.p = 1e3 * np.stack([np.roll(np.sin(2 * np.pi * 1e6 * t), k) for k in range(8)])
        sen = kSensor(m.copy())
        sen.record = ["p"]
        cases.append((kg, kWaveMedium(sound_speed=c, density=1000.0), src, sen, (8, 8), (2.0, 2.0)))
    bs = BatchSimulation(cases)
    assert len(bs.plans) == 3
    batched = bs.run()
    bs.close()
    for case, got in zip(cases, batched):
        kg, med, src, sen, pml, alpha = case
        one = Simulation(kg, med, copy.deepcopy(src), sen, pml_size=pml, pml_alpha=alpha).run()
        assert np.array_equal(got["p"], one["p"])
