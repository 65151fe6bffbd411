"""Host-side set-up of the drop-in boundary (no GPU): record parsing, validation,
source scaling, sensor tables and the device polynomial coefficients."""
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def _grid(shape, nt=10):
    from kwave_amd.kgrid import kWaveGrid

    kg = kWaveGrid(list(shape), [1e-4] * len(shape))
    kg.makeTime(1500, 0.3)
    kg.setTime(nt, float(kg.dt))
    return kg


def test_record_parsing_matches_reference_rules():
    """kspace_solver.py:281-316: shorthand expansion and dependencies."""
    from kwave_amd.solvers.setup import parse_record

    req, rec = parse_record(None, 2)
    assert req == {"p", "p_final"}
    req, rec = parse_record(["u", "p_max", "I_avg"], 3)
    assert {"ux", "uy", "uz", "p_max", "Ix_avg", "Iy_avg", "Iz_avg"} == req
    assert {"p", "ux", "uy", "uz"} <= rec
    req, rec = parse_record("u_final", 2)
    assert req == {"ux_final", "uy_final"}


def test_record_start_index_validation():
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource
    from kwave_amd.solvers.setup import build_instance

    kg = _grid((32, 32), nt=20)
    src = kSource()
    src.p0 = np.ones((32, 32))
    sen = kSensor(np.ones((32, 32), dtype=bool))
    sen.record_start_index = 5
    st = build_instance(kg, kWaveMedium(sound_speed=1500), src, sen, pml_size=(4, 4), pml_alpha=(2, 2))
    assert st.rec_start == 4 and st.Nt - st.rec_start == 16
    sen.record_start_index = 0
    with pytest.raises(ValueError, match=">= 1"):
        build_instance(kg, kWaveMedium(sound_speed=1500), src, sen, pml_size=(4, 4), pml_alpha=(2, 2))
    sen.record_start_index = 21
    with pytest.raises(ValueError, match="exceeds Nt"):
        build_instance(kg, kWaveMedium(sound_speed=1500), src, sen, pml_size=(4, 4), pml_alpha=(2, 2))


def test_bad_sensor_mask_raises():
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource
    from kwave_amd.solvers.setup import build_instance

    kg = _grid((32, 32))
    src = kSource()
    src.p0 = np.ones((32, 32))
    with pytest.raises(ValueError, match="neither binary"):
        build_instance(kg, kWaveMedium(sound_speed=1500), src, kSensor(np.ones((5, 7))), pml_size=(4, 4), pml_alpha=(2, 2))


def test_source_scaling_per_axis():
    """kspace_solver.py:612-645: dirichlet 1/(D c0^2); additive 2 dt/(D c0 d_i) (scale * fac_i); u additive 2 c0 dt / d_i."""
    from kwave_amd.kgrid import kWaveGrid
    from kwave_amd.kmedium import kWaveMedium
    from kwave_amd.ksensor import kSensor
    from kwave_amd.ksource import kSource
    from kwave_amd.solvers.setup import build_instance

    kg = kWaveGrid([32, 48], [1e-4, 2e-4])
    kg.setTime(10, 1e-8)
    c = np.full((32, 48), 1500.0)
    c[5, 7] = 1700.0
    m = np.zeros((32, 48), dtype=bool)
    m[5, 7] = m[9, 3] = True
    for mode in ("dirichlet", "additive", "additive-no-correction"):
        src = kSource()
        src.p_mask = m
        src.p = np.ones(4)
        src.p_mode = mode
        src.u_mask = m
        src.ux = np.ones(4)
        src.u_mode = mode
        st = build_instance(kg, kWaveMedium(sound_speed=c), src, kSensor(m), pml_size=(2, 2), pml_alpha=(2, 2))
        cs = np.array([1700.0, 1500.0])
        if mode == "dirichlet":
            assert np.allclose(st.src_p.scale, 1 / (2 * cs ** 2)) and st.src_p.fac[:2] == (1.0, 1.0)
            assert np.allclose(st.src_u[0].scale, 1.0)
        else:
            eff = np.outer(st.src_p.scale, st.src_p.fac[:2])
            assert np.allclose(eff, 2 * 1e-8 / (2 * cs[:, None] * np.array([1e-4, 2e-4])[None, :]))
            assert np.allclose(st.src_u[0].scale, 2 * cs * 1e-8 / 1e-4)
        assert st.src_u[1] is None


def _coeffs(src, fn):
    body = src[src.index(f"__device__ __forceinline__ float {fn}(float x2)"):]
    body = body[: body.index("\n}\n")]
    return [eval(m.group(1).replace("f", "")) for m in re.finditer(r"(?:p = |fmaf\(p, x2, )(-?[0-9.]+f?(?: / [0-9.]+f)?)", body)]


@pytest.mark.parametrize("fn,ref", [("sinc_k2", lambda x: np.sinc(x / np.pi)), ("cos_k2", np.cos)])
def test_device_kappa_polynomials(fn, ref):
    """The device evaluates kappa = sin(x)/x and kappa_s = cos(x) from x^2 by Taylor polynomials in
    x^2 on |x| <= 2; emulated in float32 they match float64 to 2e-7."""
    src = (ROOT / "k-wave-python_amd" / "csrc" / "kw_kernels.hpp").read_text()
    c = _coeffs(src, fn)
    x = np.linspace(-2, 2, 20001).astype(np.float32)
    x2 = (x * x).astype(np.float32)
    p = np.full_like(x, np.float32(c[0]))
    for cc in c[1:]:
        p = (p * x2 + np.float32(cc)).astype(np.float32)
    assert np.abs(p - ref(x.astype(np.float64))).max() < 2e-7


def test_options_and_native_bridge_defaults():
    """native.py:9-18 defaults (PML 20 / alpha 2 per axis); integer pml_size broadcast (documented deviation)."""
    from kwave_amd.options.simulation_execution_options import SimulationExecutionOptions
    from kwave_amd.options.simulation_options import SimulationOptions
    from kwave_amd.solvers.native import _per_axis

    assert _per_axis(None, 3, 20, int) == (20, 20, 20)
    assert _per_axis(SimulationOptions(pml_size=12).pml_size, 2, 20, int) == (12, 12)
    assert _per_axis([10, 14], 2, 20, int) == (10, 14)
    ex = SimulationExecutionOptions(backend="python")
    assert ex.is_python_backend
    with pytest.raises(ValueError):
        SimulationExecutionOptions(backend="cuda")


def test_unified_api_argument_errors():
    from kwave_amd.kspaceFirstOrder import _normalize_pml, _resolve_dtype

    assert _normalize_pml(10, 3) == (10, 10, 10)
    assert _normalize_pml((10, 12), 3) == (10, 12, 12)
    with pytest.raises(ValueError):
        _normalize_pml("big", 2)
    assert _resolve_dtype("single") is np.float32 and _resolve_dtype(None) is np.float64
    with pytest.raises(ValueError):
        _resolve_dtype(np.int32)


def test_plan_refuses_mixed_alpha_power():
    """The power-law exponent is a per-plan constant: a plan of absorbing instances
    with different alpha_power is refused before any device call."""
    from kwave_amd import workloads
    from kwave_amd.solvers.plan import DevicePlan
    from kwave_amd.solvers.setup import build_instance

    sts = []
    for y in (1.1, 1.5):
        kg, med, src, sen, pml, alpha = workloads.hetero_2d(seed=40, n=64, nt=20, pml=8)
        med.alpha_power = y
        sts.append(build_instance(kg, med, src, sen, pml_size=pml, pml_alpha=alpha))
    with pytest.raises(ValueError, match="alpha_power"):
        DevicePlan(sts)
