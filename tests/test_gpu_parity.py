"""HIP solver vs the FP64 oracle on seeded inputs, through the drop-in entry points.

Bar (north star): relative L2 of the recorded sensor data <= 1e-4 against the
FP64 restatement of kwave/solvers/kspace_solver.py; max-abs/peak is reported
too.  Every physics and source/sensor branch of Simulation.step
(kspace_solver.py:691-747) is exercised on grids the oracle finishes in
seconds.
"""
import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL_L2 = 1e-4


def _run_gpu(kgrid, medium, source, sensor, pml_size, pml_alpha, **kw):
    from kwave_amd.solvers.kspace_solver import Simulation

    return Simulation(kgrid, medium, copy.deepcopy(source), sensor, pml_size=pml_size, pml_alpha=pml_alpha, **kw).run()


def _run_oracle(kgrid, medium, source, sensor, pml_size, pml_alpha, **kw):
    from oracle.kspace_ref import OracleSimulation

    kw.pop("dtype", None)
    return OracleSimulation(kgrid, medium, copy.deepcopy(source), sensor, pml_size=pml_size, pml_alpha=pml_alpha, **kw).run()


def _compare(got, want, keys=None, bar=REL_L2):
    keys = keys or sorted(want)
    assert set(got) == set(want), (sorted(got), sorted(want))
    out = {}
    for k in keys:
        g, w = np.asarray(got[k], dtype=np.float64), np.asarray(want[k], dtype=np.float64)
        assert g.shape == w.shape, (k, g.shape, w.shape)
        rel = np.linalg.norm(g - w) / max(np.linalg.norm(w), 1e-300)
        mx = np.max(np.abs(g - w)) / max(np.max(np.abs(w)), 1e-300)
        out[k] = (rel, mx)
        print(f"{k}: rel-L2 {rel:.3e} max-abs/peak {mx:.3e}")
        assert rel < bar, (k, rel, mx)
    return out


def _both(case, **kw):
    args = case
    return _run_gpu(*args, **kw), _run_oracle(*args, **kw)


def _grid(n, dx=1e-4, c=1500.0, nt=120):
    from kwave_amd.data import Vector
    from kwave_amd.kgrid import kWaveGrid

    kg = kWaveGrid(Vector(list(n)), Vector([dx] * len(n)))
    kg.setTime(nt, 0.3 * dx / c)
    return kg


def _hetero(shape, seed, c_rng=(1400, 1700), r_rng=(900, 1200)):
    rng = np.random.default_rng(seed)
    from scipy.ndimage import gaussian_filter

    c = gaussian_filter(rng.uniform(*c_rng, size=shape), 2)
    r = gaussian_filter(rng.uniform(*r_rng, size=shape), 2)
    return c, r


def _p0(shape, seed):
    rng = np.random.default_rng(seed)
    p0 = np.zeros(shape)
    idx = tuple(rng.integers(s // 3, 2 * s // 3) for s in shape)
    p0[idx] = 1.0
    sl = tuple(slice(i - 2, i + 3) for i in idx)
    p0[sl] = 0.5
    return p0


# ---------------------------------------------------------------------------


def test_ivp_2d_cartesian_sensor():
    from kwave_amd import workloads

    case = workloads.ivp_2d(seed=11, n=64, pml=10)
    g, o = _both(case)
    _compare(g, o)


def test_hetero_absorbing_line_source_binary_sensor():
    from kwave_amd import workloads

    case = workloads.hetero_2d(seed=12, n=64, nt=240, pml=10)
    g, o = _both(case)
    _compare(g, o)


@pytest.mark.parametrize("shape", [(64, 64), (48, 80)])
@pytest.mark.parametrize("physics", ["lossless", "stokes", "powerlaw", "no_dispersion", "nonlinear_powerlaw"])
def test_physics_branches_ivp(shape, physics):
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource

    kg = _grid(shape, nt=100)
    c, r = _hetero(shape, 3)
    kw = {}
    if physics == "stokes":
        kw = dict(alpha_coeff=0.75, alpha_power=2.0)
    elif physics == "powerlaw":
        kw = dict(alpha_coeff=0.75, alpha_power=1.5)
    elif physics == "no_dispersion":
        kw = dict(alpha_coeff=0.75, alpha_power=1.3, alpha_mode="no_dispersion")
    elif physics == "nonlinear_powerlaw":
        kw = dict(alpha_coeff=0.5, alpha_power=1.1, BonA=6.0)
    med = kWaveMedium(sound_speed=c, density=r, **kw)
    src = kSource()
    src.p0 = 1e5 * _p0(shape, 4)
    mask = np.zeros(shape, dtype=bool)
    mask[shape[0] // 4: 3 * shape[0] // 4: 3, shape[1] // 5::4] = True
    sen = kSensor(mask)
    sen.record = ["p", "p_final", "p_max", "p_min", "p_rms"]
    g, o = _both((kg, med, src, sen, (8, 8), (2.0, 2.0)))
    _compare(g, o)


@pytest.mark.parametrize("mode", ["additive", "dirichlet", "additive-no-correction"])
@pytest.mark.parametrize("rows", ["single", "many"])
def test_pressure_sources(mode, rows):
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource

    shape = (64, 64)
    kg = _grid(shape, nt=150)
    c, r = _hetero(shape, 5)
    med = kWaveMedium(sound_speed=c, density=r)
    src = kSource()
    mask = np.zeros(shape, dtype=bool)
    mask[20, 24:40] = True
    mask[30:34, 10] = True
    src.p_mask = mask
    npt = int(mask.sum())
    t = np.arange(80) * float(kg.dt)
    base = np.sin(2 * np.pi * 1e6 * t) * np.hanning(80)
    if rows == "single":
        src.p = 1e3 * base[None, :]
    else:
        src.p = 1e3 * np.stack([np.roll(base, k) for k in range(npt)])
    src.p_mode = mode
    sen = kSensor(np.ones(shape, dtype=bool))
    sen.record = ["p", "p_final"]
    g, o = _both((kg, med, src, sen, (8, 8), (2.0, 2.0)))
    _compare(g, o)


@pytest.mark.parametrize("mode", ["additive", "dirichlet", "additive-no-correction"])
@pytest.mark.parametrize("rows", ["single", "many"])
def test_velocity_sources(mode, rows):
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource

    shape = (48, 64)
    kg = _grid(shape, nt=120)
    med = kWaveMedium(sound_speed=1500.0, density=1000.0)
    src = kSource()
    mask = np.zeros(shape, dtype=bool)
    mask[16, 20:30] = True
    src.u_mask = mask
    npt = int(mask.sum())
    t = np.arange(60) * float(kg.dt)
    base = np.sin(2 * np.pi * 1.5e6 * t) * np.hanning(60)
    if rows == "single":
        src.ux = 1e-2 * base[None, :]
        src.uy = -5e-3 * base[None, :]
    else:
        src.ux = 1e-2 * np.stack([np.roll(base, k) for k in range(npt)])
        src.uy = -5e-3 * np.stack([np.roll(base, 2 * k) for k in range(npt)])
    src.u_mode = mode
    sen = kSensor(np.ones(shape, dtype=bool))
    sen.record = ["p", "ux_staggered", "uy_staggered", "ux_final"]
    g, o = _both((kg, med, src, sen, (8, 8), (2.0, 2.0)))
    _compare(g, o)


@pytest.mark.parametrize("use_sg,use_kspace", [(False, True), (True, False), (False, False)])
def test_operator_switches(use_sg, use_kspace):
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource

    shape = (64, 48)
    kg = _grid(shape, nt=80)
    c, r = _hetero(shape, 7)
    med = kWaveMedium(sound_speed=c, density=r)
    src = kSource()
    src.p0 = _p0(shape, 8)
    sen = kSensor(np.ones(shape, dtype=bool))
    g, o = _both((kg, med, src, sen, (6, 6), (2.0, 2.0)), use_sg=use_sg, use_kspace=use_kspace)
    _compare(g, o)


def test_record_start_index_and_smoothing_off():
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource

    shape = (64, 64)
    kg = _grid(shape, nt=90)
    med = kWaveMedium(sound_speed=1500.0)
    src = kSource()
    src.p0 = _p0(shape, 9)
    mask = np.zeros(shape, dtype=bool)
    mask[40, 10:50] = True
    sen = kSensor(mask)
    sen.record = ["p", "p_max"]
    sen.record_start_index = 17
    g, o = _both((kg, med, src, sen, (10, 10), (2.0, 2.0)), smooth_p0=False)
    assert g["p"].shape == (40, 90 - 16)
    _compare(g, o)


def test_3d_nonlinear_absorbing_planar():
    from kwave_amd import workloads

    case = workloads.nonlinear_3d(seed=13, n=32, nt=60, pml=4)
    g, o = _both(case)
    _compare(g, o)


def test_3d_mixed_radix_ivp_velocity_record():
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource

    shape = (24, 40, 48)
    kg = _grid(shape, nt=50)
    c, r = _hetero(shape, 21)
    med = kWaveMedium(sound_speed=c, density=r, alpha_coeff=0.3, alpha_power=1.7)
    src = kSource()
    src.p0 = _p0(shape, 22)
    mask = np.zeros(shape, dtype=bool)
    mask[12, 5:35:3, 4:44:5] = True
    sen = kSensor(mask)
    sen.record = ["p", "uz_staggered", "ux_final", "p_final"]
    g, o = _both((kg, med, src, sen, (4, 6, 8), (2.0, 1.5, 2.5)))
    _compare(g, o)


def test_step_api_matches_oracle_state():
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource
    from kwave_amd.solvers.kspace_solver import Simulation
    from oracle.kspace_ref import OracleSimulation

    shape = (64, 64)
    kg = _grid(shape, nt=40)
    c, r = _hetero(shape, 31)
    med = kWaveMedium(sound_speed=c, density=r, BonA=4.0)
    src = kSource()
    src.p0 = 1e5 * _p0(shape, 32)
    sen = kSensor(np.ones(shape, dtype=bool))
    sim = Simulation(kg, med, copy.deepcopy(src), sen, pml_size=(10, 10), pml_alpha=(2.0, 2.0)).setup()
    ora = OracleSimulation(kg, med, copy.deepcopy(src), sen, pml_size=(10, 10), pml_alpha=(2.0, 2.0))
    for _ in range(13):
        sim.step()
        ora.step()
    assert sim.t == ora.t == 13
    for got, want in [(sim.p, ora.p)] + list(zip(sim.u, ora.u)) + list(zip(sim.rho_split, ora.rho)):
        rel = np.linalg.norm(got - want) / np.linalg.norm(want)
        assert rel < REL_L2, rel


def test_legacy_entry_point_2d():
    from kwave_amd import workloads
    from kwave_amd.kspaceFirstOrder2D import kspaceFirstOrder2D
    from kwave_amd.options.simulation_execution_options import SimulationExecutionOptions
    from kwave_amd.options.simulation_options import SimulationOptions

    kg, med, src, sen, pml, alpha = workloads.ivp_2d(seed=14, n=64, pml=10)
    opts = SimulationOptions(pml_size=[10, 10], pml_alpha=2.0)
    ex = SimulationExecutionOptions(is_gpu_simulation=True, backend="python")
    with pytest.warns(FutureWarning):
        g = kspaceFirstOrder2D(kg, copy.deepcopy(src), sen, med, opts, ex)
    o = _run_oracle(kg, med, src, sen, pml, alpha)
    _compare(g, o)


def test_unified_api_pml_outside():
    from kwave_amd.kspaceFirstOrder import kspaceFirstOrder
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource

    kg = _grid((44, 28), nt=60)
    src = kSource()
    src.p0 = _p0((44, 28), 41)
    sen = kSensor(np.ones((44, 28), dtype=bool))
    sen.record = ["p", "p_final", "p_max"]
    g = kspaceFirstOrder(kg, kWaveMedium(sound_speed=1500.0), copy.deepcopy(src), sen, pml_size=10, dtype=np.float32)
    assert g["p"].shape == (44 * 28, 60) and g["p_final"].shape == (44, 28) and g["p"].dtype == np.float32
    # same problem with the PML grown by hand and run through the oracle
    from kwave_amd.kspaceFirstOrder import _expand_for_pml_outside
    from kwave_amd.utils.filters import smooth

    kg2, med2, src2, sen2 = _expand_for_pml_outside(kg, kWaveMedium(sound_speed=1500.0), src, sen, (10, 10))
    src2.p0 = smooth(src2.p0, restore_max=True)
    o = _run_oracle(kg2, med2, src2, sen2, (10, 10), (2.0, 2.0), smooth_p0=False)
    _compare(g, o)


@pytest.mark.parametrize("shape", [(64, 48), (24, 32, 40)])
def test_velocity_and_intensity_records(shape):
    """Collocated velocity (the unstagger FFT pair, kspace_solver.py:741-743), its aggregates,
    finals and the intensity post-processing (kspace_solver.py:840-862)."""
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource

    kg = _grid(shape, nt=70)
    c, r = _hetero(shape, 51)
    med = kWaveMedium(sound_speed=c, density=r, alpha_coeff=0.4, alpha_power=1.4)
    src = kSource()
    src.p0 = _p0(shape, 52)
    mask = np.zeros(shape, dtype=bool)
    mask[tuple(slice(s // 4, 3 * s // 4, 3) for s in shape)] = True
    sen = kSensor(mask)
    ax = "xyz"[: len(shape)]
    sen.record = ["p", "u", "p_final", "u_final", "ux_max", f"u{ax[-1]}_rms", "I", "I_avg"]
    pml = tuple([6] * len(shape))
    g, o = _both((kg, med, src, sen, pml, tuple([2.0] * len(shape))))
    _compare(g, o)


def test_velocity_record_with_time_varying_source():
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource

    shape = (64, 64)
    kg = _grid(shape, nt=120)
    med = kWaveMedium(sound_speed=1500.0, density=1000.0)
    src = kSource()
    m = np.zeros(shape, dtype=bool)
    m[20, 20:44] = True
    src.p_mask = m
    t = np.arange(60) * float(kg.dt)
    src.p = 1e3 * np.sin(2 * np.pi * 1e6 * t)[None, :] * np.hanning(60)[None, :]
    sen = kSensor(np.ones(shape, dtype=bool))
    sen.record = ["p", "ux", "uy", "uy_min"]
    g, o = _both((kg, med, src, sen, (8, 8), (2.0, 2.0)))
    _compare(g, o)


def test_time_reversal_reconstruction():
    """TimeReversal (reconstruction/time_reversal.py) on the GPU vs the same steps on the oracle."""
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource
    from kwave_amd.kspaceFirstOrder2D import kspaceFirstOrder2D
    from kwave_amd.options.simulation_execution_options import SimulationExecutionOptions
    from kwave_amd.options.simulation_options import SimulationOptions
    from kwave_amd.reconstruction import TimeReversal
    from oracle.kspace_ref import OracleSimulation

    shape = (64, 64)
    kg = _grid(shape, nt=150)
    med = kWaveMedium(sound_speed=1500.0)
    src = kSource()
    src.p0 = _p0(shape, 61)
    mask = np.zeros(shape, dtype=bool)
    mask[10, 10:54] = mask[53, 10:54] = mask[10:54, 10] = mask[10:54, 53] = True
    sen = kSensor(mask)
    sen.record = ["p"]
    opts = SimulationOptions(pml_size=[8, 8], pml_alpha=2.0)
    ex = SimulationExecutionOptions(is_gpu_simulation=True, backend="python")
    import warnings

    with warnings.catch_warnings():
        warnings.simplefilter("ignore", FutureWarning)
        fwd = kspaceFirstOrder2D(kg, copy.deepcopy(src), sen, med, opts, ex)
        sen.recorded_pressure = fwd["p"]
        tr = TimeReversal(kg, med, sen)
        recon = tr(kspaceFirstOrder2D, opts, ex)
    rsrc = kSource()
    rsrc.p_mask = mask
    rsrc.p = np.flip(fwd["p"], axis=1)
    rsrc.p_mode = "dirichlet"
    want = OracleSimulation(kg, med, rsrc, kSensor(mask, record=["p", "p_final"]), pml_size=(8, 8), pml_alpha=(2.0, 2.0)).run()
    want_recon = np.clip(2.0 * want["p_final"], 0, None).T
    assert recon.shape == want_recon.shape
    rel = np.linalg.norm(recon - want_recon) / np.linalg.norm(want_recon)
    print("time reversal rel-L2", rel)
    assert rel < 1e-4
