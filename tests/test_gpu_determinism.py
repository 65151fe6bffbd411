"""Run-to-run determinism of the HIP solver.

Two runs of the same simulation must give bit-identical outputs: the solver
has no atomics and a fixed reduction order, so any difference means a data
race (for example between LDS phases of consecutive transforms in one
workgroup).  The shapes give workgroups of two waves (length 80) and four
waves (length 512, as in the 512^3 benchmark) on the power-law + nonlinear
path, which runs two forward transforms back to back.
"""
import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(shape, nt=80):
    from kwave_amd.data import Vector
    from kwave_amd.kgrid import kWaveGrid
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource

    rng = np.random.default_rng(21)
    dx = 1e-4
    kg = kWaveGrid(Vector(list(shape)), Vector([dx] * len(shape)))
    kg.setTime(nt, 0.3 * dx / 1700.0)
    c = 1500.0 + 200.0 * rng.random(shape)
    med = kWaveMedium(sound_speed=c, density=1000.0 + 100.0 * rng.random(shape), alpha_coeff=0.75, alpha_power=1.5, BonA=6.0)
    src = kSource()
    p0 = np.zeros(shape)
    p0[tuple(s // 2 for s in shape)] = 1e5
    src.p0 = p0
    mask = np.zeros(shape, dtype=bool)
    mask[(slice(None),) * (len(shape) - 1) + (shape[-1] // 3,)] = True
    sen = kSensor(mask)
    sen.record = ["p", "p_final"]
    return kg, med, src, sen


@pytest.mark.parametrize("shape", [(48, 80), (16, 512), (24, 32, 40)])
def test_repeat_runs_bit_identical(shape):
    from kwave_amd.solvers.kspace_solver import Simulation

    kg, med, src, sen = _case(shape)
    pml = tuple(4 for _ in shape)
    outs = [Simulation(kg, med, copy.deepcopy(src), sen, pml_size=pml, pml_alpha=tuple(2.0 for _ in shape)).run() for _ in range(3)]
    for k in outs[0]:
        for o in outs[1:]:
            assert np.array_equal(outs[0][k], o[k]), (shape, k, float(np.max(np.abs(outs[0][k] - o[k]))))
    assert np.all(np.isfinite(outs[0]["p"]))


@pytest.mark.parametrize("shape", [(64, 80), (24, 32, 40)])
def test_column_tile_variants_bit_identical(shape, monkeypatch):
    """Both compiled tile widths of the column passes give the same bits (each
    column is transformed by the same code whatever its neighbours), so the
    per-plan tile autotuning cannot change results."""
    from kwave_amd.solvers.kspace_solver import Simulation

    kg, med, src, sen = _case(shape)
    pml = tuple(4 for _ in shape)
    outs = []
    for tiles in ("00", "11"):
        monkeypatch.setenv("KW_COL_TILES", tiles)
        outs.append(Simulation(kg, med, copy.deepcopy(src), sen, pml_size=pml, pml_alpha=tuple(2.0 for _ in shape)).run())
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), (shape, k)


@pytest.mark.parametrize("shape", [(64, 80), (24, 32, 40), (40, 512, 24), (512, 24, 40)])
@pytest.mark.parametrize("grid", ["8", "16", "24"])
def test_looping_column_passes_bit_identical(shape, grid, monkeypatch):
    """Column passes whose workgroups loop over tiles (and load the next tile's
    input during the current tile) give the bits of one workgroup per tile.  The
    benchmark grids loop only at full size; KW_COL_GRID forces a few workgroups
    per instance here, so every workgroup runs several tiles, with ragged last
    tiles and the 512-point kernels of the benchmark (KW_COL_WIDE keeps their
    large-grid tile widths)."""
    from kwave_amd.solvers.kspace_solver import Simulation

    kg, med, src, sen = _case(shape, nt=40)
    pml = tuple(4 for _ in shape)
    outs = []
    monkeypatch.setenv("KW_COL_WIDE", "1")
    for persist, g in (("0", "0"), ("3", grid)):
        monkeypatch.setenv("KW_COL_PERSIST", persist)
        monkeypatch.setenv("KW_COL_GRID", g)
        outs.append(Simulation(kg, med, copy.deepcopy(src), sen, pml_size=pml, pml_alpha=tuple(2.0 for _ in shape)).run())
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), (shape, grid, k)


@pytest.mark.parametrize("shape", [(80, 112), (40, 56, 48), (320, 64), (448, 64)])
def test_radix_plan_variants_agree_with_the_oracle(shape, monkeypatch):
    """Lengths with a factor 5 or 7 have two radix plans for the column passes (radix 2 or
    radix 4 next to the odd factor); the plan picks one by timing, KW_COL_PLANS forces it.
    Both transform the same lines to FP32 rounding (the two runs differ by rounding alone,
    about 1e-5 after 50 nonlinear power-law steps), and both match the FP64 oracle.  The
    sensor line sits 6 cells from the point source along the last axis, inside the ~13
    cells the pulse covers in 50 steps: a line beyond that reach records only the
    spectral tails of the point source, ~1e-6 of its peak, whose relative error is
    rounding noise rather than a property of the transforms."""
    from test_gpu_parity import _compare, _run_oracle

    from kwave_amd.solvers.kspace_solver import Simulation

    kg, med, src, sen = _case(shape, nt=50)
    mask = np.zeros(shape, dtype=bool)
    mask[(slice(None),) * (len(shape) - 1) + (shape[-1] // 2 - 6,)] = True
    sen.mask = mask
    pml = tuple(4 for _ in shape)
    alpha = tuple(2.0 for _ in shape)
    outs = []
    for plans in ("00", "11"):
        monkeypatch.setenv("KW_COL_PLANS", plans)
        outs.append(Simulation(kg, med, copy.deepcopy(src), sen, pml_size=pml, pml_alpha=alpha, smooth_p0=False).run())
    for k in outs[0]:
        rel = float(np.linalg.norm(outs[0][k] - outs[1][k]) / max(np.linalg.norm(outs[0][k]), 1e-30))
        assert rel < 5e-5, (shape, k, rel)
    want = _run_oracle(kg, med, copy.deepcopy(src), sen, pml, alpha, smooth_p0=False)
    _compare(outs[0], want)
    _compare(outs[1], want)


def _source_case(shape, per_point, nt=83, slen=29):
    """Multi-row additive p and u sources whose signals end before nt: one
    signal row (the dense-pattern form) or one row per point (the spectral
    form)."""
    from kwave_amd.data import Vector
    from kwave_amd.kgrid import kWaveGrid
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource

    rng = np.random.default_rng(5)
    dx = 1e-4
    kg = kWaveGrid(Vector(list(shape)), Vector([dx] * len(shape)))
    kg.setTime(nt, 0.3 * dx / 1700.0)
    med = kWaveMedium(sound_speed=1500.0 + 200.0 * rng.random(shape), density=1000.0, alpha_coeff=0.5, alpha_power=1.3)
    src = kSource()
    pm = np.zeros(shape, dtype=bool)
    pm[(slice(shape[0] // 4, shape[0] // 4 + 6),) + tuple(slice(s // 2 - 2, s // 2 + 2) for s in shape[1:])] = True
    src.p_mask = pm
    um = np.zeros(shape, dtype=bool)
    um[(slice(shape[0] // 2, shape[0] // 2 + 3),) + tuple(slice(s // 3, s // 3 + 3) for s in shape[1:])] = True
    src.u_mask = um
    np_, nu = (int(pm.sum()), int(um.sum())) if per_point else (1, 1)
    src.p = 1e4 * np.sin(np.arange(slen)[None, :] * 0.4 + np.arange(np_)[:, None] * 0.1)
    src.ux = 0.05 * np.cos(np.arange(slen - 7)[None, :] * 0.3 + np.arange(nu)[:, None] * 0.2)
    sm = np.zeros(shape, dtype=bool)
    sm[(slice(None),) * (len(shape) - 1) + (shape[-1] // 3,)] = True
    sen = kSensor(sm)
    sen.record = ["p", "u", "p_final"]
    return kg, med, src, sen


@pytest.mark.parametrize("per_point", [False, True], ids=["pattern", "spectral"])
@pytest.mark.parametrize("shape", [(64, 80), (24, 32, 40)])
def test_lean_steps_after_sources_end_bit_identical(shape, per_point, monkeypatch):
    """Once every filtered source has ended the step leaves out the source
    transforms and the staged source term; results equal the full step's bit
    for bit, through run() (graph blocks split at the switch), step() and the
    graph-free path."""
    from kwave_amd.solvers.kspace_solver import Simulation

    kg, med, src, sen = _source_case(shape, per_point)
    pml = tuple(4 for _ in shape)
    alpha = tuple(2.0 for _ in shape)

    def run(env, stepwise=False):
        for k in ("KW_NO_LEAN", "KW_NO_GRAPH"):
            monkeypatch.delenv(k, raising=False)
        for k in env:
            monkeypatch.setenv(k, "1")
        sim = Simulation(kg, med, copy.deepcopy(src), sen, pml_size=pml, pml_alpha=alpha)
        if not stepwise:
            return sim.run()
        sim.setup()
        while sim.t < sim.Nt:
            sim.step()
        return sim.run()

    ref = run(("KW_NO_LEAN",))
    assert np.abs(ref["p"]).max() > 0
    for out in (run(()), run((), stepwise=True), run(("KW_NO_GRAPH",))):
        for k in ref:
            assert np.array_equal(ref[k], out[k]), (shape, k)


def test_lean_steps_in_a_batch_wait_for_the_longest_source(monkeypatch):
    """A batch switches to the lean form only after the last of its instances'
    filtered sources has ended (a dense-pattern source of 29 samples beside a
    spectral one of 50); every instance keeps the full form's bits."""
    from kwave_amd.batch import BatchSimulation

    shape = (64, 80)
    pml, alpha = (4, 4), (2.0, 2.0)
    cases = [(*_source_case(shape, False, slen=29), pml, alpha), (*_source_case(shape, True, slen=50), pml, alpha)]
    outs = []
    for lean in (False, True):
        monkeypatch.delenv("KW_NO_LEAN", raising=False)
        if not lean:
            monkeypatch.setenv("KW_NO_LEAN", "1")
        bs = BatchSimulation([(kg, med, copy.deepcopy(src), sen, p, a) for kg, med, src, sen, p, a in cases])
        assert len(bs.plans) == 1
        try:
            outs.append(bs.run())
        finally:
            bs.close()
    for full, lean in zip(*outs):
        for k in full:
            assert np.array_equal(full[k], lean[k]), k
