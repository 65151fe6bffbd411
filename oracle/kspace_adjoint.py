"""FP64 discrete adjoint of the k-space time loop (gradient with respect to sound speed).

TEST INFRASTRUCTURE ONLY, like ``oracle/kspace_ref.py``: it is the checker of the
HIP adjoint; only ``tests/`` import it.

The reference has no adjoint; this differentiates its forward step
(kwave/solvers/kspace_solver.py:691-747, restated in OracleSimulation.step) with a
binary sensor; first the linear lossless case with an additive pressure source:

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

Absorbing media (kspace_solver.py:471-516, linear: no B/A) add to the equation of
state c^2 (tau A - eta B) with A = L1[rho0 d] (Stokes: rho0 d), d = sum_i Dop_i u_i,
B = L2[sum_i rho_i]; L1, L2 are real k-space multipliers (self-adjoint).  With
w = c^2 pbar the reverse step gains
    cbar += w ((y-1) tau A - y eta B) / c          (d tau/dc, d eta/dc at fixed A, B)
    rbar_i += w - L2[eta w]                         (instead of += w)
    ubar_i += Gop_i(b_i dt rho0 rbar_i - rho0 L1[tau w])   (Dop_i^T = -Gop_i)
with A and B of each step kept from the forward sweep.

Nonlinear media (B/A, kspace_solver.py:537-547): the equation of state gains
c^2 B/A S^2 / (2 rho0), S = sum_i rho_i, and the mass update multiplies Dop_i u_i by
(2 S^{t-1} + rho0) / rho0.  The reverse step then takes w (1 + B/A S^t / rho0) into
rbar_i, multiplies the gradient-job input b_i dt rho0 rbar_i by that factor, and adds
-2 dt sum_i b_i d_i^t rbar_i (d_i^t = Dop_i u_i^t kept from the forward sweep) to every
rbar_j of step t-1.

Dirichlet pressure sources write sig / (D c^2) into every rho_i at their points after
the mass update: cbar gains -2 v / c sum_i rbar_i there and rbar_i is cleared.
Velocity sources (kspace_solver.py:628-645, 707) act on u_i^t after the momentum
update: additive ones add K[scale sig] with scale = 2 c dt / d_i, so cbar[pts] gains
(scale / c) sig(t) [K ubar_i][pts] with ubar_i = dJ/du_i^t; dirichlet ones write sig
(no c-dependence) and clear ubar_i at the points.

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
    mask = np.array(mask_raw, dtype=bool).ravel()
    if mask.size == 1:
        mask = np.full(sim.c0.size, bool(mask[0]))
    pts = np.where(mask)[0]
    sig = np.array(sig_raw, dtype=np.float64)
    sig = sig.reshape(1, -1) if sig.ndim == 1 else sig.reshape(-1, sig.shape[-1])
    rows = np.zeros(pts.size, dtype=int) if sig.shape[0] == 1 else np.arange(pts.size)
    c = sim.c0.ravel()[pts]
    # additive: scale_i = base / d_i (kspace_solver.py:619-624); dirichlet: the value written
    # into every rho_i is sig base (:613-617)
    base = 1.0 / (sim.ndim * c ** 2) if mode == "dirichlet" else 2 * sim.dt / (sim.ndim * c)
    return pts, sig, rows, base, c, mode


def _velocity_source_points(sim):
    """Per axis: (points, signal, signal rows, scale, c, mode) of the velocity source, or None.
    additive: u_i += K[scale sig] with scale = 2 c dt / d_i; dirichlet: u_i = sig at the
    points (kspace_solver.py:628-645)."""
    src = sim.source
    mask_raw = getattr(src, "u_mask", 0)
    mode = getattr(src, "u_mode", None) or "additive"
    out = []
    for i, name in enumerate(["ux", "uy", "uz"][: sim.ndim]):
        sig_raw = getattr(src, name, 0)
        if not (_is_enabled(mask_raw) and _is_enabled(sig_raw)):
            out.append(None)
            continue
        mask = np.array(mask_raw, dtype=bool).ravel()
        if mask.size == 1:
            mask = np.full(sim.c0.size, bool(mask[0]))
        pts = np.where(mask)[0]
        sig = np.array(sig_raw, dtype=np.float64)
        sig = sig.reshape(1, -1) if sig.ndim == 1 else sig.reshape(-1, sig.shape[-1])
        rows = np.zeros(pts.size, dtype=int) if sig.shape[0] == 1 else np.arange(pts.size)
        c = sim.c0.ravel()[pts] if np.ndim(sim.c0) else np.full(pts.size, float(sim.c0))
        scale = np.ones_like(c) if mode == "dirichlet" else 2 * c * sim.dt / sim.spacing[i]
        out.append((pts, sig, rows, scale, c, mode))
    return out if any(o is not None for o in out) else None


def _velocity_source_term(sim, ubar_i, term, t):
    """Adjoint of step t's velocity source on axis i, ubar_i = dJ/du_i^t: (ubar_i, the dJ/dc
    increments at the points).  additive: d(scale)/dc = scale / c against K ubar_i
    (uncorrected: ubar_i); dirichlet: the written value does not depend on c, and the
    overwritten velocity carries no adjoint further back."""
    pts, sig, rows, scale, c, mode = term
    if mode == "dirichlet":
        f = ubar_i.ravel().copy()
        f[pts] = 0.0
        return f.reshape(ubar_i.shape), 0.0
    Z = sim._diff(ubar_i, sim.source_kappa) if mode == "additive" else ubar_i
    return ubar_i, (scale / c) * sig[rows, t] * Z.ravel()[pts]


def sound_speed_gradient(sim: OracleSimulation, adj_src):
    """Run ``sim`` forward from t = 0 and return (dJ/dc0 as a grid, recorded p).

    ``adj_src`` is dJ/d(recorded p), shape (n_sensor, n_rec) like ``run()["p"]``.
    """
    if sim.t != 0:
        raise ValueError("adjoint: pass a fresh simulation")
    if sim.cart is not None or sim.sensor_idx is None:
        raise ValueError("adjoint: binary sensor masks only")
    nd, shape = sim.ndim, sim.shape
    adj_src = np.asarray(adj_src, dtype=np.float64)
    absorbing = sim.tau is not None or sim.eta is not None
    nonlinear = sim.BonA is not None
    ps, As, Bs, Ss, Ds = [], [], [], [], []
    while sim.t < sim.Nt:
        sim.step()
        ps.append(sim.p.copy())
        if absorbing:  # the absorption and dispersion operands of this step's equation of state
            As.append(_absorption_operand(sim))
            Bs.append(sim._diff(sum(sim.rho), sim.nabla2) if sim.eta is not None else None)
        if nonlinear:  # sum of the densities after the step, and each axis's Dop_i u_i
            Ss.append(sum(sim.rho))
            Ds.append([sim._diff(sim.u[i], sim.op_div[i]) for i in range(nd)])
    rec = sim.data["p"].copy()
    src = _source_points(sim)
    usrc = _velocity_source_points(sim)
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
        w = c0_sq * pbar
        cbar += 2 * ps[t] * pbar / c0
        dbar = 0.0
        wc = w
        if absorbing:
            cbar += _absorption_speed_term(sim, c0, w, As[t], Bs[t])
            dbar, w = _absorption_adjoint(sim, rho0, w)
        if nonlinear:
            w = w + _nonlinear_eos_term(sim, rho0, wc, Ss[t])
        rbar = [rbar[i] + w for i in range(nd)]
        if src is not None:
            pts, sig, rows, base, c, mode = src
            if t < sig.shape[1] and mode == "dirichlet":
                rbar, dc = _dirichlet_source_term(rbar, pts, base * sig[rows, t], c)
                flat = cbar.ravel()
                flat[pts] += dc
                cbar = flat.reshape(shape)
            elif t < sig.shape[1]:
                R = sum(rbar[i] / sim.spacing[i] for i in range(nd))
                Z = sim._diff(R, sim.source_kappa) if mode == "additive" else R
                flat = cbar.ravel()
                flat[pts] += -(base / c) * sig[rows, t] * Z.ravel()[pts]
                cbar = flat.reshape(shape)
        nl = 1.0
        sbar = 0.0
        if nonlinear:  # the mass update of step t multiplies by (2 S^{t-1} + rho0) / rho0
            s_prev = Ss[t - 1] if t > 0 else 0.0
            nl = (2.0 * s_prev + rho0) / rho0
            sbar = _nonlinear_mass_term(sim, rbar, Ds[t])
        for i in range(nd):
            w = sim.pml[i] * sim.dt * rho0 * nl * rbar[i] - dbar
            rbar[i] = sim.pml[i] ** 2 * rbar[i] + sbar
            ubar[i] = ubar[i] + sim._diff(w, sim.op_grad[i])
        if usrc is not None:  # ubar_i is now dJ/du_i^t, after step t's velocity source
            for i, term in enumerate(usrc):
                if term is not None and t < term[1].shape[1]:
                    ubar[i], dc = _velocity_source_term(sim, ubar[i], term, t)
                    flat = cbar.ravel()
                    flat[term[0]] += dc
                    cbar = flat.reshape(shape)
        pnew = np.zeros(shape)
        for i in range(nd):
            v = sim.pml_sg[i] * sim.dt_rho0[i] * ubar[i]
            ubar[i] = sim.pml_sg[i] ** 2 * ubar[i]
            pnew += sim._diff(v, sim.op_div[i])
        pbar = pnew
    if sim.p0 is not None:
        cbar += _initial_density_term(sim, c0, rbar)
    return cbar, rec


def _absorption_operand(sim):
    """What tau multiplies in the equation of state of the step just taken: rho0 div u
    (Stokes) or L1[rho0 div u] (power law), div u = sum_i Dop_i u_i (kspace_solver.py:712-724)."""
    if sim.tau is None:
        return None
    div = sum(sim._diff(sim.u[i], sim.op_div[i]) for i in range(sim.ndim))
    return sim.rho0 * div if sim.stokes else sim._diff(sim.rho0 * div, sim.nabla1)


def _absorption_adjoint(sim, rho0, w):
    """(adjoint of div u through the absorption term, what rbar_i gains) for w = c^2 pbar:
    rho0 L1[tau w] (Stokes: rho0 tau w), and w - L2[eta w] (the dispersion term is
    -eta L2[sum rho])."""
    dbar = 0.0
    if sim.tau is not None:
        dbar = rho0 * (sim.tau * w if sim.stokes else sim._diff(sim.tau * w, sim.nabla1))
    if sim.eta is not None:
        w = w - sim._diff(sim.eta * w, sim.nabla2)
    return dbar, w


def _absorption_speed_term(sim, c0, w, A, B):
    """c-derivative of c^2 (tau A - eta B) at fixed A, B against pbar (w = c^2 pbar):
    tau = -2 a c^(y-1) and eta = 2 a c^y tan(pi y / 2) give d tau/dc = (y-1) tau / c and
    d eta/dc = y eta / c (Stokes: tau = -2 a c, y = 2)."""
    y = sim.alpha_power
    out = 0.0
    if sim.tau is not None:
        out = out + w * (y - 1.0) * sim.tau / c0 * A
    if sim.eta is not None:
        out = out - w * y * sim.eta / c0 * B
    return out


def _nonlinear_eos_term(sim, rho0, w, S):
    """What each rbar_i gains from the B/A term c^2 B/A S^2 / (2 rho0) of the equation of
    state (kspace_solver.py:546): w B/A S / rho0, w = c^2 pbar."""
    return w * sim.BonA * S / rho0


def _nonlinear_mass_term(sim, rbar, D):
    """Adjoint of S^{t-1} through the factor (2 S^{t-1} + rho0) / rho0 of the mass update
    rho_i^t = b_i (b_i rho_i^{t-1} - dt rho0 d_i nl) (kspace_solver.py:547, :713-719):
    -2 dt sum_i b_i d_i rbar_i, added to every rbar_j of step t-1."""
    return -2.0 * sim.dt * sum(sim.pml[i] * D[i] * rbar[i] for i in range(sim.ndim))


def _dirichlet_source_term(rbar, pts, v, c):
    """A dirichlet pressure source writes v = sig / (D c^2) into every rho_i at its points
    after the mass update (kspace_solver.py:86-91, :613-617): dJ/dc gains -2 v / c sum_i
    rbar_i there, and the overwritten values carry no adjoint further back.  Returns
    (rbar with the points cleared, the dJ/dc increments at the points)."""
    dc = -2.0 * v / c * sum(r.ravel()[pts] for r in rbar)
    out = []
    for r in rbar:
        f = r.ravel().copy()
        f[pts] = 0.0
        out.append(f.reshape(r.shape))
    return out, dc


def _initial_density_term(sim, c0, rbar):
    """d/dc of rho_i^0 = p0 / (D c^2) against the adjoints rbar_i of rho_i^0 (kspace_solver.py:732)."""
    return -2.0 * sim.p0 / (sim.ndim * c0 ** 3) * sum(rbar)
