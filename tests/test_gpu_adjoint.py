"""Reverse-time adjoint on the GPU (kw_plan_adjoint_c) against the FP64 adjoint and finite differences.

Functional J = 1/2 sum (y(c) - d)^2 over the recorded pressure traces; the
adjoint source is dJ/dy = y - d.  Checks:

* the GPU gradient equals the FP64 CPU adjoint (oracle/kspace_adjoint.py, itself
  pinned by finite differences in tests/test_oracle_adjoint.py) to FP32 accuracy,
  in 2-D with all three pressure-source modes and in 3-D, with an initial pressure p0
  beside the line source and with p0 alone, in power-law (with and without
  dispersion) and Stokes media, with B/A nonlinearity (alone, and with p0 and
  power-law absorption), with velocity sources in each mode, on grids of the
  generic-length family, and in 1-D;
* central finite differences of the GPU solver at 20 seeded cells (seed 7) of the
  128^2 surrogate problem, 600 steps, in the region the wave reaches and returns
  from.  h = 5 m/s: the FP32 solver's rounding noise in J (~1e4 of J ~ 1e13 over
  19200 trace samples) divided by 2h swamps the derivative at h = 0.1 m/s, while the
  truncation error of central differences at h/c = 0.3 % stays below 1e-5;
* the result does not depend on the checkpoint interval (bit-identical);
* a batch of instances gives each instance's single-plan gradient (bit-identical);
* torch.autograd through kwave_amd.adjoint.sensor_traces gives the same gradient.
"""
import copy

import numpy as np
import pytest

from test_oracle_adjoint import ABSORBING, NONLINEAR, adjoint_case

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def _gpu_grad(case, adj, checkpoint=50, c=None):
    from kwave_amd.adjoint import sound_speed_gradient

    kg, med, src, sen, pml, alpha = case
    return sound_speed_gradient(kg, med, src, sen, adj, c=c, pml_size=pml, pml_alpha=alpha, checkpoint=checkpoint)


def _gpu_traces(case, c=None):
    from kwave_amd.adjoint import forward_traces

    kg, med, src, sen, pml, alpha = case
    return forward_traces(kg, med, src, sen, c=c, pml_size=pml, pml_alpha=alpha)


def _oracle_grad(case, adj):
    from oracle.kspace_adjoint import sound_speed_gradient
    from oracle.kspace_ref import OracleSimulation

    kg, med, src, sen, pml, alpha = case
    sim = OracleSimulation(kg, med, copy.deepcopy(src), sen, pml_size=pml, pml_alpha=alpha)
    return sound_speed_gradient(sim, adj)


def _misfit_data(y, seed=5):
    return 0.7 * y + 0.05 * np.abs(y).max() * np.random.default_rng(seed).standard_normal(y.shape)


@pytest.mark.parametrize("shape,mode,p0,line,alpha", [
    ((32, 40), "additive", False, True, None), ((32, 40), "additive-no-correction", False, True, None),
    ((16, 24, 32), "additive", False, True, None), ((32, 40), "additive", True, True, None),
    ((32, 40), None, True, False, None), ((16, 24, 32), None, True, False, None),
    ((32, 40), "additive", False, True, "powerlaw"), ((32, 40), "additive", False, True, "powerlaw-no-dispersion"),
    ((32, 40), "additive", True, True, "stokes"), ((16, 24, 32), "additive", False, True, "powerlaw-mild"),
    ((32, 40), "dirichlet", False, True, None), ((16, 24, 32), "dirichlet", True, True, "powerlaw-mild")])
def test_gpu_adjoint_matches_fp64_adjoint(shape, mode, p0, line, alpha):
    """Line source (additive, uncorrected or dirichlet), initial pressure p0 beside it, or p0
    alone (the initial-density term of dJ/dc0); lossless, power-law (with and without
    dispersion) and Stokes media.  The 3-D
    power-law case absorbs less: at 40 dB/(MHz^1.5 cm) the 16-cell grid keeps 17 % of the
    lossless amplitude and the FP32 forward traces sit 8.5e-4 from FP64."""
    case = adjoint_case(shape=shape, nt=90 if len(shape) == 2 else 60, mode=mode or "additive", p0=p0, line_source=line,
                        alpha=ABSORBING.get(alpha))
    y = _gpu_traces(case)
    d = _misfit_data(y)
    g_gpu, y2 = _gpu_grad(case, y - d, checkpoint=16)
    assert np.array_equal(y, y2)
    g_ref, y_ref = _oracle_grad(case, y - d)
    assert _rel(y, y_ref) < 1e-4
    err = _rel(g_gpu, g_ref)
    print(f"{shape} {mode} p0={p0} {alpha}: gradient rel-L2 vs FP64 adjoint {err:.2e}")
    assert err < 1e-3, err


@pytest.mark.parametrize("shape,p0,line,alpha", [
    ((32, 40), False, True, None), ((32, 40), True, True, "powerlaw-mild"), ((16, 24, 32), True, False, None),
    ((32, 40), False, True, "stokes")])
def test_gpu_adjoint_nonlinear_matches_fp64_adjoint(shape, p0, line, alpha):
    """B/A = 6 at a source amplitude of 30 MPa (B/A S / rho0 reaches a few per cent near the
    source): the equation of state's B/A term and the mass update's (2 S + rho0) / rho0 factor."""
    case = adjoint_case(shape=shape, nt=90 if len(shape) == 2 else 60, p0=p0, line_source=line, alpha=ABSORBING.get(alpha),
                        **NONLINEAR)
    y = _gpu_traces(case)
    d = _misfit_data(y)
    g_gpu, y2 = _gpu_grad(case, y - d, checkpoint=16)
    assert np.array_equal(y, y2)
    g_ref, y_ref = _oracle_grad(case, y - d)
    assert _rel(y, y_ref) < 1e-4
    err = _rel(g_gpu, g_ref)
    print(f"{shape} B/A p0={p0} {alpha}: gradient rel-L2 vs FP64 adjoint {err:.2e}")
    assert err < 1e-3, err


@pytest.mark.parametrize("shape,umode,p0,alpha,nonlin", [
    ((32, 40), "additive", False, None, False), ((32, 40), "additive-no-correction", False, None, False),
    ((32, 40), "dirichlet", False, None, False), ((16, 24, 32), "additive", True, "powerlaw-mild", True)])
def test_gpu_adjoint_velocity_source_matches_fp64_adjoint(shape, umode, p0, alpha, nonlin):
    """A velocity source on every axis beside the pressure line source: the speed term of the
    additive source scale 2 c dt / d_i (K-filtered, or not), and a dirichlet source's
    cleared points."""
    case = adjoint_case(shape=shape, nt=90 if len(shape) == 2 else 60, p0=p0, alpha=ABSORBING.get(alpha), u_source=umode,
                        **(NONLINEAR if nonlin else {}))
    y = _gpu_traces(case)
    d = _misfit_data(y)
    g_gpu, y2 = _gpu_grad(case, y - d, checkpoint=16)
    assert np.array_equal(y, y2)
    g_ref, y_ref = _oracle_grad(case, y - d)
    assert _rel(y, y_ref) < 1e-4
    err = _rel(g_gpu, g_ref)
    print(f"{shape} u:{umode} p0={p0} {alpha} B/A={nonlin}: gradient rel-L2 vs FP64 adjoint {err:.2e}")
    assert err < 1e-3, err


@pytest.mark.parametrize("shape,mode,p0,alpha,nonlin,umode", [
    ((30, 44), "additive", False, None, False, None), ((33, 40), "dirichlet", True, "powerlaw", False, None),
    ((30, 44), "additive", True, "powerlaw-mild", True, "additive"), ((18, 20, 22), "additive", False, "stokes", True, "dirichlet")])
def test_gpu_adjoint_generic_lengths_match_fp64_adjoint(shape, mode, p0, alpha, nonlin, umode):
    """Grids that run on the generic-length family: axis lengths with a prime factor above 7
    (30 = 2 3 5 across the columns, 44 = 4 11 and 22 along the contiguous axis) or an odd
    row count (33 rows), so the reverse step's contiguous-axis work is unfused and the
    column passes use the stash spectrum."""
    case = adjoint_case(shape=shape, nt=90 if len(shape) == 2 else 60, mode=mode, p0=p0, alpha=ABSORBING.get(alpha), u_source=umode,
                        **(NONLINEAR if nonlin else {}))
    y = _gpu_traces(case)
    d = _misfit_data(y)
    g_gpu, y2 = _gpu_grad(case, y - d, checkpoint=16)
    assert np.array_equal(y, y2)
    g_ref, y_ref = _oracle_grad(case, y - d)
    assert _rel(y, y_ref) < 1e-4
    err = _rel(g_gpu, g_ref)
    print(f"{shape} {mode} p0={p0} {alpha} B/A={nonlin} u:{umode}: gradient rel-L2 vs FP64 adjoint {err:.2e}")
    assert err < 1e-3, err


@pytest.mark.parametrize("mode,amp", [("additive", None), ("dirichlet", None), ("additive", 3e7)])
def test_gpu_adjoint_1d_matches_fp64_adjoint(mode, amp):
    """1-D (32 points, run as the 2 x N embedding, both rows' gradients summed): a point
    pressure source beside p0, power law and an additive velocity source, linear or with
    B/A = 6 at a source amplitude amp.  The sensor sits 15 points from the source, as in
    the 2-D cases: on 64 points it would lie beyond the ~27 points a pulse covers in 90
    steps and record only the sources' spectral tails, whose relative error is FP32
    rounding noise (1.8e-4 in the traces there)."""
    nl = dict(bona=6.0, amplitude=amp) if amp else {}
    case = adjoint_case(shape=(32,), mode=mode, p0=True, alpha=ABSORBING["powerlaw"], u_source="additive", **nl)
    y = _gpu_traces(case)
    d = _misfit_data(y)
    g_gpu, y2 = _gpu_grad(case, y - d, checkpoint=16)
    assert np.array_equal(y, y2)
    assert g_gpu.shape == (32,)
    g_ref, y_ref = _oracle_grad(case, y - d)
    fwd = _rel(y, y_ref)
    err = _rel(g_gpu, g_ref)
    print(f"1-D {mode} B/A amplitude {amp}: traces rel-L2 {fwd:.2e}, gradient rel-L2 vs FP64 adjoint {err:.2e}")
    assert fwd < 1e-4
    assert err < 1e-3, err


def test_gpu_adjoint_cartesian_sensor_is_the_transpose_of_its_interpolation():
    """Off-grid Cartesian sensor points record W p, W the linear-interpolation weights of
    their corner nodes: the traces equal W times the traces of a binary mask over those
    corners, and the gradient for an adjoint source r equals the binary-mask gradient for
    W^T r (the scatter of the reverse sweep is the gather's transpose; its atomic adds of
    shared corners change the rounding only)."""
    from kwave_amd.ksensor import kSensor
    from kwave_amd.solvers.setup import linear_interp_table

    kg, med, src, sen, pml, alpha = adjoint_case()
    shape, dx = (32, 40), 1e-4
    rng = np.random.default_rng(11)
    pts = np.stack([(rng.uniform(20.0, 24.0, 12) - shape[0] // 2) * dx, (rng.uniform(8.0, 30.0, 12) - shape[1] // 2) * dx])
    idx, w = linear_interp_table(pts, shape, (dx, dx))
    corners = np.unique(idx.ravel())  # a binary mask records in ascending flat index
    W = np.zeros((pts.shape[1], corners.size))
    for p in range(pts.shape[1]):
        for c in range(idx.shape[1]):
            W[p, np.searchsorted(corners, idx[p, c])] += w[p, c]
    sen_c = kSensor(pts)
    sen_c.record = ["p"]
    bm = np.zeros(int(np.prod(shape)), dtype=bool)
    bm[corners] = True
    sen_b = kSensor(bm.reshape(shape))
    sen_b.record = ["p"]
    case_c = (kg, med, src, sen_c, pml, alpha)
    case_b = (kg, med, src, sen_b, pml, alpha)
    yc, yb = _gpu_traces(case_c), _gpu_traces(case_b)
    assert _rel(yc, W @ yb) < 1e-6
    r = yc - _misfit_data(yc)
    gc, _ = _gpu_grad(case_c, r, checkpoint=16)
    gb, _ = _gpu_grad(case_b, (W.T @ r).astype(np.float32), checkpoint=16)
    err = _rel(gc, gb)
    print(f"Cartesian vs binary-corner gradient: rel-L2 {err:.2e}")
    assert err < 1e-5, err


def test_gpu_adjoint_matches_finite_differences_surrogate_problem():
    from kwave_amd import workloads

    case = workloads.surrogate_2d(seed=5_000_123)
    c = np.asarray(case[1].sound_speed, dtype=np.float64)
    y = _gpu_traces(case).astype(np.float64)
    d = _misfit_data(y)
    grad, _ = _gpu_grad(case, (y - d).astype(np.float32))
    rng = np.random.default_rng(7)
    ii, jj = np.meshgrid(np.arange(c.shape[0]), np.arange(c.shape[1]), indexing="ij")
    near = (ii >= 22) & (ii < 64) & (jj >= 36) & (jj < 92) & (c < c.max() - 10.0)
    cells = rng.choice(np.where(near.ravel())[0], size=20, replace=False)
    h = 5.0

    def J(cc):
        return 0.5 * float(np.sum((_gpu_traces(case, c=cc).astype(np.float64) - d) ** 2))

    fd, ad = [], []
    for cell in cells:
        cp, cm = c.copy(), c.copy()
        cp.ravel()[cell] += h
        cm.ravel()[cell] -= h
        fd.append((J(cp) - J(cm)) / (2 * h))
        ad.append(grad.ravel()[cell])
    fd, ad = np.array(fd), np.array(ad)
    err = _rel(ad, fd)
    print(f"adjoint vs central differences at 20 cells: rel-L2 {err:.2e}; max |g| {np.abs(ad).max():.3e}")
    assert err < 2e-2, (err, ad, fd)


def test_gpu_adjoint_checkpoint_interval_and_batch():
    from kwave_amd import _abi
    from kwave_amd.solvers.plan import DevicePlan
    from kwave_amd.solvers.setup import build_instance

    cases = [adjoint_case(seed=s) for s in (7, 8)]
    sts = [build_instance(kg, med, copy.deepcopy(src), sen, pml_size=pml, pml_alpha=al) for kg, med, src, sen, pml, al in cases]
    singles = []
    for st in sts:
        pl = DevicePlan([st])
        pl.run(st.Nt)
        y = pl.recorded(0, _abi.F_REC_P).copy()
        adj = (y - _misfit_data(y)).astype(np.float32)
        g7 = pl.adjoint_c([adj], checkpoint=7)[0]
        g50 = pl.adjoint_c([adj], checkpoint=50)[0]
        assert np.array_equal(g7, g50)
        singles.append((adj, g7))
        pl.close()
    pl = DevicePlan(sts)
    gb = pl.adjoint_c([a for a, _ in singles], checkpoint=25)
    pl.close()
    for (a, g), gbi in zip(singles, gb):
        assert np.array_equal(g, gbi)


def test_torch_autograd_matches_adjoint():
    import torch

    from kwave_amd.adjoint import sensor_traces

    kg, med, src, sen, pml, alpha = case = adjoint_case()
    c = torch.tensor(np.asarray(med.sound_speed, dtype=np.float64), requires_grad=True)
    y = sensor_traces(c, kg, med, src, sen, pml_size=pml, pml_alpha=alpha, checkpoint=30)
    d = torch.from_numpy(_misfit_data(y.detach().numpy()).astype(np.float32))
    loss = 0.5 * ((y - d) ** 2).sum()
    loss.backward()
    g_ref, _ = _gpu_grad(case, (y.detach() - d).numpy(), checkpoint=30)
    assert np.array_equal(c.grad.numpy(), g_ref)
