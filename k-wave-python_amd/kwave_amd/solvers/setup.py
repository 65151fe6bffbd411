"""FP64 host-side construction of everything one solver instance needs.

This is the set-up half of the reference ``Simulation`` (kwave/solvers/
kspace_solver.py:205-689): grid, medium expansion, c_ref, sensor parsing,
PML profiles, absorption/dispersion/nonlinearity coefficients, source scaling,
staggered density and p0 smoothing.  Everything is computed in float64 and
handed to the device as float32; the time loop itself runs in HIP kernels.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from ..utils.conversion import db2neper
from ..utils.filters import smooth
from ..utils.pml import get_pml


def is_enabled(x):
    """Zero arrays/scalars mean "disabled" (kspace_solver.py:23-24)."""
    return x is not None and not (np.all(np.asarray(x) == 0) if hasattr(x, "__len__") else x == 0)


def expand_to_grid(val, shape, name):
    """Scalar -> full grid, or reshape a grid-sized array (kspace_solver.py:48-57)."""
    if val is None:
        raise ValueError(f"Missing required parameter: {name}")
    arr = np.array(val, dtype=np.float64).ravel()
    n = int(np.prod(shape))
    if arr.size == 1:
        return np.full(shape, arr[0])
    if arr.size == n:
        return arr.reshape(shape)
    raise ValueError(f"{name} size {arr.size} incompatible with grid size {n}")


def scalar_or_grid(val, shape, name):
    """(value, is_scalar): keep homogeneous properties scalar so the device reads one float."""
    arr = np.array(val, dtype=np.float64).ravel()
    if arr.size == 1:
        return arr[0], True
    return expand_to_grid(val, shape, name), False


def stagger(arr, axis):
    """Midpoint average along ``axis`` keeping the last plane (kspace_solver.py:798-809)."""
    lo = [slice(None)] * arr.ndim
    hi = [slice(None)] * arr.ndim
    lo[axis], hi[axis] = slice(None, -1), slice(1, None)
    last = [slice(None)] * arr.ndim
    last[axis] = slice(-1, None)
    return np.concatenate([0.5 * (arr[tuple(lo)] + arr[tuple(hi)]), arr[tuple(last)]], axis=axis)


def grid_shape_and_spacing(kgrid):
    """kWaveGrid (.N/.spacing) or a namespace with Nx/dx, Ny/dy, Nz/dz (kspace_solver.py:209-223)."""
    if hasattr(kgrid, "N") and hasattr(kgrid, "spacing"):
        return tuple(int(n) for n in np.atleast_1d(kgrid.N)), [float(d) for d in np.atleast_1d(kgrid.spacing)]
    shape, spacing = [], []
    for a in "xyz":
        N, d = getattr(kgrid, f"N{a}", None), getattr(kgrid, f"d{a}", None)
        if N is None or d is None:
            break
        shape.append(int(N))
        spacing.append(float(d))
    return tuple(shape), spacing


def parse_record(record, ndim):
    """Expand record shorthands and add dependencies (kspace_solver.py:281-316).

    Returns (requested, internal) sets of keys.
    """
    rec = record or ("p", "p_final")
    if isinstance(rec, str):
        rec = (rec,)
    rec = set(rec)
    ax = "xyz"[:ndim]
    if "u" in rec:
        rec.discard("u")
        rec.update(f"u{a}" for a in ax)
    if "u_staggered" in rec:
        rec.discard("u_staggered")
        rec.update(f"u{a}_staggered" for a in ax)
    for suf in ("_max", "_min", "_rms", "_final"):
        if f"u{suf}" in rec:
            rec.discard(f"u{suf}")
            rec.update(f"u{a}{suf}" for a in ax)
    if "I" in rec:
        rec.discard("I")
        rec.update(f"I{a}" for a in ax)
    if "I_avg" in rec:
        rec.discard("I_avg")
        rec.update(f"I{a}_avg" for a in ax)
    requested = set(rec)
    for k in list(rec):
        for suf in ("_max", "_min", "_rms"):
            if k.endswith(suf):
                rec.add(k[: -len(suf)])
    if any(k.startswith("I") for k in requested):
        rec.update(["p"] + [f"u{a}" for a in ax])
    return requested, rec


def linear_interp_table(cart, shape, spacing):
    """Indices and weights reproducing RegularGridInterpolator(method="linear", bounds_error=True).

    Axis coordinates are (arange(N) - N//2) * d (kspace_solver.py:341); each
    query point gets 2^D (flat C-order index, weight) pairs.
    """
    nd = len(shape)
    cart = np.asarray(cart, dtype=np.float64)
    npts = cart.shape[1]
    lo_idx, frac = [], []
    for a in range(nd):
        x = (np.arange(shape[a]) - shape[a] // 2) * spacing[a]
        q = cart[a]
        if np.any(q < x[0]) or np.any(q > x[-1]) or not np.all(np.isfinite(q)):
            raise ValueError(f"One of the requested xi is out of bounds in dimension {a}")
        i = np.searchsorted(x, q) - 1
        i = np.clip(i, 0, shape[a] - 2)
        lo_idx.append(i)
        frac.append((q - x[i]) / (x[i + 1] - x[i]))
    ncorner = 2 ** nd
    idx = np.zeros((npts, ncorner), dtype=np.int64)
    w = np.ones((npts, ncorner), dtype=np.float64)
    strides = [int(np.prod(shape[a + 1:])) for a in range(nd)]
    for c in range(ncorner):
        for a in range(nd):
            bit = (c >> (nd - 1 - a)) & 1
            idx[:, c] += (lo_idx[a] + bit) * strides[a]
            w[:, c] *= frac[a] if bit else (1.0 - frac[a])
    return idx, w


@dataclass
class SourceTerm:
    mode: int              # _abi.SRC_*
    idx: np.ndarray        # ascending C-order flat indices (int64)
    scale: np.ndarray      # per point (float64)
    sigrow: np.ndarray     # per point (int32)
    sig: np.ndarray        # [nsigrow, len] float64
    fac: tuple             # per-axis multiplier


@dataclass
class InstanceSetup:
    """Everything the device needs for one simulation, in float64."""
    shape: tuple
    spacing: list
    ndim: int
    Nt: int
    dt: float
    c_ref: float
    use_sg: bool
    use_kspace: bool
    pml: list
    pml_sg: list
    pml_sizes: list
    c0: np.ndarray
    c0_is_scalar: bool
    rho0: np.ndarray
    rho0_is_scalar: bool
    dtr0: list
    flags: int = 0
    alpha_power: float = 1.5
    tau: object = None
    eta: object = None
    alpha_np: float = None             # scalar alpha: tau/eta derived from c0 on the device
    bona: object = None
    src_p: SourceTerm = None
    src_u: list = field(default_factory=list)
    p0: np.ndarray = None
    sensor_idx: np.ndarray = None      # [nsens, ncorner] int64
    sensor_w: np.ndarray = None        # [nsens, ncorner] or None
    n_sensor: int = 0
    requested: set = field(default_factory=set)
    record: set = field(default_factory=set)
    rec_start: int = 0
    p0_split: tuple = None             # share of p0 / c0^2 per split density at t = 0 (None: 1/ndim each)


def build_instance(kgrid, medium, source, sensor, *, use_sg=True, use_kspace=True, smooth_p0=True,
                   pml_size=None, pml_alpha=None, source_scaled=False, rho0_sg=None):
    """Host set-up of one instance (kspace_solver.py:205-689), float64 throughout.

    ``source_scaled``: the time-varying source signals are already in the units the
    time loop injects (the k-Wave binaries' input files hold them so): no
    2 dt/(D c0 d) / 1/(D c0^2) / 2 c0 dt/d scaling, and a pressure source adds
    (or sets) its value in every split density.  ``rho0_sg``: staggered-grid
    densities per axis (scalars or grids) to use instead of the midpoint average.
    """
    from .. import _abi

    shape, spacing = grid_shape_and_spacing(kgrid)
    nd = len(shape)
    Nt, dt = int(kgrid.Nt), float(kgrid.dt)
    c_val, c_scalar = scalar_or_grid(medium.sound_speed, shape, "sound_speed")
    dens = getattr(medium, "density", None)
    r_val, r_scalar = scalar_or_grid(dens if dens is not None else 1000.0, shape, "density")
    c_ref = float(np.max(c_val))

    # --- sensor (kspace_solver.py:243-326) ---
    numel = int(np.prod(shape))
    mask = getattr(sensor, "mask", None) if sensor is not None else None
    sw = None
    if mask is None:
        sidx = np.arange(numel, dtype=np.int64)[:, None]
    else:
        m = np.asarray(mask, dtype=float)
        if m.ndim == 2 and m.shape[0] == nd:
            sidx, sw = linear_interp_table(m, shape, spacing)
        elif m.size == 1 or m.size == numel:
            bm = np.asarray(mask, dtype=bool).ravel()
            if bm.size == 1:
                bm = np.full(numel, bool(bm[0]))
            sidx = np.where(bm)[0].astype(np.int64)[:, None]
        else:
            raise ValueError(f"Sensor mask shape {m.shape} is neither binary (numel={numel}) nor Cartesian ({nd}, N_points)")
    requested, record = parse_record(getattr(sensor, "record", None) if sensor is not None else None, nd)
    raw = getattr(sensor, "record_start_index", None) if sensor is not None else None
    start = int(raw) if raw is not None else 1
    if start < 1:
        raise ValueError(f"sensor.record_start_index must be >= 1 (1-based, MATLAB convention), got {start}")
    if start > Nt:
        raise ValueError(f"sensor.record_start_index ({start}) exceeds Nt ({Nt})")

    # --- PML (kspace_solver.py:363-406) ---
    pml, pml_sg, pml_sizes = [], [], []
    for ax in range(nd):
        name = "xyz"[ax]
        size = int(pml_size[ax]) if pml_size is not None else int(getattr(kgrid, f"pml_size_{name}", 0))
        alpha = float(pml_alpha[ax]) if pml_alpha is not None else float(getattr(kgrid, f"pml_alpha_{name}", 0))
        pml_sizes.append(size if alpha != 0 else 0)
        if size == 0 or alpha == 0:
            pml.append(np.ones(shape[ax]))
            pml_sg.append(np.ones(shape[ax]))
        else:
            pml.append(get_pml(shape[ax], spacing[ax], dt, c_ref, size, alpha, staggered=False).ravel())
            pml_sg.append(get_pml(shape[ax], spacing[ax], dt, c_ref, size, alpha, staggered=True).ravel())

    # --- staggered density (kspace_solver.py:656-663) ---
    if rho0_sg is not None:
        dtr0 = []
        for a in range(nd):
            sg = np.asarray(rho0_sg[a], dtype=np.float64)
            dtr0.append(dt / (float(sg.ravel()[0]) if sg.size == 1 else expand_to_grid(sg, shape, "rho0_sg")))
    elif r_scalar:
        dtr0 = [dt / r_val] * nd
    else:
        dtr0 = [dt / (stagger(r_val, a) if use_sg else r_val) for a in range(nd)]

    st = InstanceSetup(shape=shape, spacing=spacing, ndim=nd, Nt=Nt, dt=dt, c_ref=c_ref, use_sg=use_sg,
                       use_kspace=use_kspace, pml=pml, pml_sg=pml_sg, pml_sizes=pml_sizes, c0=c_val,
                       c0_is_scalar=c_scalar, rho0=r_val, rho0_is_scalar=r_scalar, dtr0=dtr0,
                       sensor_idx=sidx, sensor_w=sw, n_sensor=sidx.shape[0], requested=requested,
                       record=record, rec_start=start - 1)

    # --- absorption, dispersion, nonlinearity (kspace_solver.py:449-550) ---
    mode = getattr(medium, "alpha_mode", None)
    mode = getattr(mode, "value", mode)
    a_raw = getattr(medium, "alpha_coeff", 0)
    if is_enabled(a_raw):
        y_raw = getattr(medium, "alpha_power", None)
        y = float(np.array(1.5 if y_raw is None else y_raw).flatten()[0])
        st.alpha_power = y
        a_val, a_scalar = scalar_or_grid(a_raw, shape, "alpha_coeff")
        a_np = db2neper(a_val, y)
        stokes = abs(y - 2.0) < 1e-10
        if a_scalar:
            st.alpha_np = float(a_np)
        if mode != "no_absorption":
            if stokes:
                st.flags |= _abi.KW_STOKES
                if not a_scalar:
                    st.tau = -2 * a_np * c_val
            else:
                st.flags |= _abi.KW_POWERLAW_ABS
                if not a_scalar:
                    st.tau = -2 * a_np * c_val ** (y - 1)
        if mode not in ("no_absorption", "no_dispersion") and not stokes:
            st.flags |= _abi.KW_POWERLAW_DISP
            if not a_scalar:
                st.eta = 2 * a_np * c_val ** y * np.tan(np.pi * y / 2)
    b_raw = getattr(medium, "BonA", 0)
    if is_enabled(b_raw):
        st.flags |= _abi.KW_NONLINEAR
        st.bona, _ = scalar_or_grid(b_raw, shape, "BonA")

    # --- sources (kspace_solver.py:552-645, 60-114) ---
    c0_flat = np.broadcast_to(c_val, shape).ravel()

    def term(mask_raw, sig_raw, mode_str, scale_fn, fac):
        mk = np.array(mask_raw, dtype=bool).ravel()
        if mk.size == 1:
            mk = np.full(numel, bool(mk[0]))
        if mk.size != numel:
            raise ValueError(f"source mask has {mk.size} elements but grid has {numel} points.")
        pts = np.where(mk)[0].astype(np.int64)
        s = np.array(sig_raw, dtype=np.float64)
        if s.ndim == 1:
            s = s.reshape(1, -1)
        elif s.ndim > 2:
            s = s.reshape(-1, s.shape[-1])
        if s.shape[0] == 1:
            sigrow = np.zeros(pts.size, dtype=np.int32)
        elif s.shape[0] == pts.size:
            sigrow = np.arange(pts.size, dtype=np.int32)
        else:
            raise ValueError(f"source signal has {s.shape[0]} rows for {pts.size} source points")
        modes = {"dirichlet": _abi.SRC_DIRICHLET, "additive": _abi.SRC_ADDITIVE, "additive-no-correction": _abi.SRC_ADDITIVE_NOCORR}
        if mode_str not in modes:
            raise ValueError(f"Unknown source mode: {mode_str!r}. Use 'additive', 'additive-no-correction', or 'dirichlet'.")
        return SourceTerm(modes[mode_str], pts, scale_fn(c0_flat[pts]), sigrow, s, tuple(fac) + (1.0,) * (3 - len(fac)))

    src = source
    p_mask, p_sig = getattr(src, "p_mask", 0), getattr(src, "p", 0)
    p_mode = getattr(src, "p_mode", None) or "additive"
    if is_enabled(p_mask) and is_enabled(p_sig) and source_scaled:
        st.src_p = term(p_mask, p_sig, p_mode, lambda c: np.ones_like(c), [1.0] * nd)
    elif is_enabled(p_mask) and is_enabled(p_sig):
        if p_mode == "dirichlet":
            st.src_p = term(p_mask, p_sig, p_mode, lambda c: 1.0 / (nd * c ** 2), [1.0] * nd)
        else:
            st.src_p = term(p_mask, p_sig, p_mode, lambda c: 2 * dt / (nd * c), [1.0 / d for d in spacing])
    u_mask = getattr(src, "u_mask", 0)
    u_mode = getattr(src, "u_mode", None) or "additive"
    st.src_u = []
    for i, name in enumerate(["ux", "uy", "uz"][:nd]):
        sig = getattr(src, name, 0)
        if is_enabled(u_mask) and is_enabled(sig):
            if u_mode == "dirichlet" or source_scaled:
                fn = lambda c: np.ones_like(c)  # noqa: E731
            else:
                fn = lambda c, _d=spacing[i]: 2 * c * dt / _d  # noqa: E731
            st.src_u.append(term(u_mask, sig, u_mode, fn, [1.0] * nd))
        else:
            st.src_u.append(None)

    # --- initial pressure (kspace_solver.py:678-689) ---
    p0_raw = getattr(src, "p0", 0)
    if is_enabled(p0_raw):
        p0 = expand_to_grid(p0_raw, shape, "p0")
        if smooth_p0 and nd >= 2:
            p0 = np.asarray(smooth(p0, restore_max=True), dtype=np.float64)
        st.p0 = p0
    return st
