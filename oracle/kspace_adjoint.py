"""FP64 discrete adjoint of the k-space time loop (gradient with respect to sound speed).

TEST INFRASTRUCTURE ONLY, like ``oracle/kspace_ref.py``: it is the checker of the
HIP adjoint; only ``tests/`` import it.

The reference has no adjoint; this differentiates its forward step
(kwave/solvers/kspace_solver.py:691-747, restated in OracleSimulation.step) for
the linear lossless case with an additive pressure source and a binary sensor:

    u_i^t = a_i^2 u_i^{t-1} - a_i beta_i Gop_i p^{t-1}          (:702-709)
    rho_i^t = b_i^2 rho_i^{t-1} - b_i dt rho0 Dop_i u_i^t + s_i(t) (:711-721)
    p^t = c^2 sum_i rho_i^t,   y_t = S p^t                        (:723-747)

with Gop_i = F^-1 op_grad_i F, Dop_i = F^-1 op_div_i F (kspace_solver.py:443-444),
a = pml_sg, b = pml, beta_i = dt / rho0_sg_i, and the additive source
s_i(t) = K M (scale_i * sig(t)), K = F^-1 kappa_s F (or identity without the
correction), scale_i = 2 dt / (D c d_i) at the mask points (:619-624).  The
operator tables are Hermitian (real parts at Nyquist), so Gop_i^T = -Dop_i,
Dop_i^T = -Gop_i and K^T = K.  For J with dJ/dy_t = r_t the reverse sweep is

    pbar += S^T r_t
    rbar_i += c^2 pbar;            cbar += 2 p^t pbar / c
    cbar[pts] += -(scale/c) sig(t) [K sum_i rbar_i / d_i][pts]    (source scale)
    ubar_i += Gop_i(b_i dt rho0 rbar_i);  rbar_i <- b_i^2 rbar_i
    pbar = sum_i Dop_i(a_i beta_i ubar_i);  ubar_i <- a_i^2 ubar_i

c_ref, dt, the PML and kappa are held fixed (they depend on c only through
c_ref = max(c0), which the gradient treats as frozen).

With an initial pressure p0 the reference overrides the state after step 0
(kspace_solver.py:729-733): p^0 = p0, rho_i^0 = p0 / (D c^2) and
u_i^0 = (dt / (2 rho0_sg_i)) Gop_i p0, so nothing of step 0 itself reaches the
traces.  Of that state only rho_i^0 depends on c, which gives the last term
    cbar += -2 p0 / (D c^3) sum_i rbar_i
with rbar_i the adjoint of rho_i^0 (what the reverse sweep of step 1 leaves);
the record of p^0 = p0 and the step-0 source term carry no c-dependence.

Pins: finite differences of OracleSimulation in FP64 (tests/test_oracle_adjoint.py).
"""
from __future__ import annotations

import numpy as np

from .kspace_ref import OracleSimulation, _is_enabled


def _source_points(sim):
    src = sim.source
    mask_raw, sig_raw = getattr(src, "p_mask", 0), getattr(src, "p", 0)
    if not (_is_enabled(mask_raw) and _is_enabled(sig_raw)):
        return None
    mode = getattr(src, "p_mode", None) or "additive"
    if mode == "dirichlet":
        raise ValueError("adjoint: dirichlet pressure sources are not differentiated")
    mask = np.array(mask_raw, dtype=bool).ravel()
    if mask.size == 1:
        mask = np.full(sim.c0.size, bool(mask[0]))
    pts = np.where(mask)[0]
    sig = np.array(sig_raw, dtype=np.float64)
    sig = sig.reshape(1, -1) if sig.ndim == 1 else sig.reshape(-1, sig.shape[-1])
    rows = np.zeros(pts.size, dtype=int) if sig.shape[0] == 1 else np.arange(pts.size)
    c = sim.c0.ravel()[pts]
    base = 2 * sim.dt / (sim.ndim * c)  # scale_i = base / d_i
    return pts, sig, rows, base, c, mode


def sound_speed_gradient(sim: OracleSimulation, adj_src):
    """Run ``sim`` forward from t = 0 and return (dJ/dc0 as a grid, recorded p).

    ``adj_src`` is dJ/d(recorded p), shape (n_sensor, n_rec) like ``run()["p"]``.
    """
    if sim.t != 0:
        raise ValueError("adjoint: pass a fresh simulation")
    if sim.tau is not None or sim.eta is not None or sim.BonA is not None:
        raise ValueError("adjoint: linear lossless media only")
    if sim.cart is not None or sim.sensor_idx is None:
        raise ValueError("adjoint: binary sensor masks only")
    if any(_is_enabled(getattr(sim.source, n, 0)) for n in ("ux", "uy", "uz")):
        raise ValueError("adjoint: velocity sources are not differentiated")
    nd, shape = sim.ndim, sim.shape
    adj_src = np.asarray(adj_src, dtype=np.float64)
    ps = []
    while sim.t < sim.Nt:
        sim.step()
        ps.append(sim.p.copy())
    rec = sim.data["p"].copy()
    src = _source_points(sim)
    c0 = sim.c0 if np.ndim(sim.c0) else np.full(shape, float(sim.c0))
    c0_sq = c0 ** 2
    rho0 = sim.rho0 if np.ndim(sim.rho0) else np.full(shape, float(sim.rho0))

    pbar = np.zeros(shape)
    ubar = [np.zeros(shape) for _ in range(nd)]
    rbar = [np.zeros(shape) for _ in range(nd)]
    cbar = np.zeros(shape)
    t_end = 1 if sim.p0 is not None else 0  # with p0 the state after step 0 is set, not stepped
    for t in range(sim.Nt - 1, t_end - 1, -1):
        if t >= sim.start:
            flat = pbar.ravel()
            flat[sim.sensor_idx] += adj_src[:, t - sim.start]
            pbar = flat.reshape(shape)
        rbar = [rbar[i] + c0_sq * pbar for i in range(nd)]
        cbar += 2 * ps[t] * pbar / c0
        if src is not None:
            pts, sig, rows, base, c, mode = src
            if t < sig.shape[1]:
                R = sum(rbar[i] / sim.spacing[i] for i in range(nd))
                Z = sim._diff(R, sim.source_kappa) if mode == "additive" else R
                flat = cbar.ravel()
                flat[pts] += -(base / c) * sig[rows, t] * Z.ravel()[pts]
                cbar = flat.reshape(shape)
        for i in range(nd):
            w = sim.pml[i] * sim.dt * rho0 * rbar[i]
            rbar[i] = sim.pml[i] ** 2 * rbar[i]
            ubar[i] = ubar[i] + sim._diff(w, sim.op_grad[i])
        pnew = np.zeros(shape)
        for i in range(nd):
            v = sim.pml_sg[i] * sim.dt_rho0[i] * ubar[i]
            ubar[i] = sim.pml_sg[i] ** 2 * ubar[i]
            pnew += sim._diff(v, sim.op_div[i])
        pbar = pnew
    if sim.p0 is not None:
        cbar += _initial_density_term(sim, c0, rbar)
    return cbar, rec


def _initial_density_term(sim, c0, rbar):
    """d/dc of rho_i^0 = p0 / (D c^2) against the adjoints rbar_i of rho_i^0 (kspace_solver.py:732)."""
    return -2.0 * sim.p0 / (sim.ndim * c0 ** 3) * sum(rbar)
