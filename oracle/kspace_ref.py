"""FP64 CPU oracle for the k-space pseudospectral time loop.

TEST INFRASTRUCTURE ONLY.  This module is the checker that the HIP solver is
compared against; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product path never
calls it.

It restates, in float64 NumPy, the algorithm of k-wave-python's native solver
``kwave/solvers/kspace_solver.py`` (class ``Simulation``, lines 122-816, and the
post-processing helpers at 824-862).  Each function cites the reference lines it
follows.  The arithmetic is deliberately op-for-op the same as the reference
(full complex ``numpy.fft.fftn/ifftn``, per-axis additive-source FFT pairs), so
that its cost also models the reference's CPU path.

Pins (see tests/test_oracle_*.py):
  * closed form: homogeneous lossless IVP with PML off equals
    ``F^-1[cos(n c|k| dt) F[p0]]`` (exactness of the k-space leapfrog);
  * physics of plane waves (tests/test_oracle_physics.py): power-law decay
    ``exp(-alpha_0 w^y x)`` for y = 1.1, 1.5 and Stokes y = 2, the
    Kramers-Kronig phase speed of kwave/utils/mapgen.py:328-358, second-
    harmonic growth of the Fubini solution with beta = 1 + B/2A (the B/A term
    and the (2 sum(rho) + rho0)/rho0 mass factor), the intensity
    ``I_avg = P^2 / (2 rho c)`` and the near-unity ``alpha_power`` stability
    case of tests/test_issue_664_alpha_power_near_unity.py, each with a planted
    mistake that the check must catch;
  * ``get_pml`` closed form and reference test values (tests/test_pmlutils.py);
  * ``db2neper`` known answer (tests/test_utils.py:31-35);
  * ``smooth``/``make_disc``/``_stagger`` through the reference's SHA-256
    checksum fixtures (tests/reference_outputs/*.json);
  * Cartesian sensors through ``scipy.interpolate.RegularGridInterpolator``;
  * linearity, dirichlet-source and mirror-symmetry invariants.
"""
from __future__ import annotations

import math

import numpy as np

__all__ = [
    "get_pml",
    "db2neper",
    "get_win_blackman_rotated",
    "smooth",
    "OracleSimulation",
    "compute_aggregates",
    "acoustic_intensity",
]


# ---------------------------------------------------------------------------
# helpers restated from kwave/utils
# ---------------------------------------------------------------------------

def get_pml(N, dx, dt, c, pml_size, pml_alpha, staggered):
    """1-D PML multiplier; kwave/utils/pml.py:6-64 (dimension-free form).

    Left:  exp(-alpha (c/dx) ((x - L - 1)/(-L))^4 dt/2) for x = 1..L
    Right: exp(-alpha (c/dx) (x/L)^4 dt/2); the staggered variant uses x + 0.5.
    """
    L = int(pml_size)
    x = np.arange(1, L + 1, dtype=np.float64)
    xs = x + 0.5 if staggered else x
    left = pml_alpha * (c / dx) * (((xs - L - 1) / (-L)) ** 4)
    right = pml_alpha * (c / dx) * ((xs / L) ** 4)
    prof = np.ones(int(N), dtype=np.float64)
    prof[:L] = np.exp(-left * dt / 2)
    prof[int(N) - L:] = np.exp(-right * dt / 2)
    return prof


def db2neper(alpha, y=1):
    """dB/(MHz^y cm) -> Np/((rad/s)^y m); kwave/utils/conversion.py:19-35."""
    return 100.0 * np.asarray(alpha, dtype=np.float64) * (1e-6 / (2.0 * math.pi)) ** y / (20.0 * np.log10(np.exp(1)))


def _blackman_1d(L, param=0.16):
    """Symmetric Blackman cosine series of length L; kwave/utils/signals.py:91-108,180-181."""
    n = np.arange(L, dtype=np.float64)
    a0, a1, a2 = (1 - param) / 2, 0.5, param / 2
    return a0 - a1 * np.cos(2 * np.pi * n / (L - 1)) + a2 * np.cos(4 * np.pi * n / (L - 1))


def get_win_blackman_rotated(shape, symmetric):
    """Rotated Blackman window used by ``smooth``; kwave/utils/signals.py:117-302.

    ``symmetric[i]`` False enlarges axis i by one before building and trims the
    last sample afterwards (signals.py:134, 246-248, 291-293).
    """
    from scipy.interpolate import interp1d

    shape = tuple(int(s) for s in shape)
    sym = np.broadcast_to(np.asarray(symmetric, dtype=bool), (len(shape),))
    Nbig = tuple(s + (0 if f else 1) for s, f in zip(shape, sym))
    L = max(Nbig)
    lin = _blackman_1d(L)
    radius = (L - 1) / 2
    ll = np.linspace(-radius, radius, L)
    axes = [np.linspace(-radius, radius, n) for n in Nbig]
    grids = np.meshgrid(*axes, indexing="ij")
    r = np.sqrt(sum(g ** 2 for g in grids))
    r[r > radius] = radius
    win = interp1d(ll, lin)(r)
    return win[tuple(slice(0, s) for s in shape)]


def smooth(a, restore_max=False):
    """Blackman-window spectral smoothing; kwave/utils/filters.py:549-615 (ndim >= 2)."""
    a = np.asarray(a, dtype=np.float64)
    orig_shape = a.shape
    sq = np.squeeze(a) if a.ndim != np.squeeze(a).ndim else a
    shape = sq.shape
    if len(shape) == 1:
        raise ValueError("oracle smooth handles ndim >= 2 only")
    sym = np.array([s % 2 == 1 for s in shape])
    win = np.abs(get_win_blackman_rotated(shape, sym))
    out = np.real(np.fft.ifftn(np.fft.fftn(sq) * np.fft.ifftshift(win))).reshape(orig_shape)
    if restore_max:
        out = (np.abs(a).max() / np.abs(out).max()) * out
    return out


def _is_enabled(x):
    """kspace_solver.py:23-24."""
    return x is not None and not (np.all(np.asarray(x) == 0) if hasattr(x, "__len__") else x == 0)


def _expand(val, shape, name):
    """kspace_solver.py:48-57."""
    if val is None:
        raise ValueError(f"Missing required parameter: {name}")
    arr = np.array(val, dtype=np.float64).ravel()
    n = int(np.prod(shape))
    if arr.size == 1:
        return np.full(shape, arr[0])
    if arr.size == n:
        return arr.reshape(shape)
    raise ValueError(f"{name} size {arr.size} incompatible with grid size {n}")


# ---------------------------------------------------------------------------
# the time loop
# ---------------------------------------------------------------------------

class OracleSimulation:
    """FP64 restatement of ``Simulation`` (kwave/solvers/kspace_solver.py:122-778).

    ``fft_workers`` > 1 runs the N-D transforms through ``scipy.fft`` with that
    many threads (the same pocketfft algorithm as numpy.fft; a multi-core timing
    variant of the CPU baseline); the default is numpy.fft, as the reference.
    """

    def __init__(self, kgrid, medium, source, sensor, *, use_sg=True, use_kspace=True,
                 smooth_p0=True, pml_size=None, pml_alpha=None, fft_workers=1):
        self.kgrid, self.medium, self.source, self.sensor = kgrid, medium, source, sensor
        self.use_sg, self.use_kspace, self.smooth_p0 = use_sg, use_kspace, smooth_p0
        self._pml_size = pml_size
        self._pml_alpha = pml_alpha
        if fft_workers > 1:
            import scipy.fft as sfft

            self._fftn = lambda a: sfft.fftn(a, workers=fft_workers)
            self._ifftn = lambda a: sfft.ifftn(a, workers=fft_workers)
        else:
            self._fftn, self._ifftn = np.fft.fftn, np.fft.ifftn
        self.t = 0
        self._setup()

    # -- setup: kspace_solver.py:205-241 --------------------------------------
    def _setup(self):
        kg = self.kgrid
        if hasattr(kg, "N") and hasattr(kg, "spacing"):
            self.shape = tuple(int(n) for n in kg.N)
            self.spacing = [float(d) for d in kg.spacing]
        else:
            shp, sp = [], []
            for a in "xyz":
                N, d = getattr(kg, f"N{a}", None), getattr(kg, f"d{a}", None)
                if N is None or d is None:
                    break
                shp.append(int(N))
                sp.append(float(d))
            self.shape, self.spacing = tuple(shp), sp
        self.ndim = len(self.shape)
        self.Nt = int(kg.Nt)
        self.dt = float(kg.dt)
        self.c0 = _expand(self.medium.sound_speed, self.shape, "sound_speed")
        dens = getattr(self.medium, "density", None)
        self.rho0 = _expand(dens if dens is not None else 1000.0, self.shape, "density")
        self.c_ref = float(np.max(self.c0))
        self._setup_sensor()
        self._setup_pml()
        self._setup_kspace()
        self._setup_physics()
        self._setup_sources()
        self._setup_fields()

    def _setup_sensor(self):
        """kspace_solver.py:243-326 (mask kinds, record parsing, start index)."""
        nd, numel = self.ndim, int(np.prod(self.shape))
        mask = getattr(self.sensor, "mask", None)
        self.cart = None
        if mask is None:
            self.n_sensor = numel
            self.sensor_idx = np.arange(numel)
        else:
            m = np.asarray(mask, dtype=float)
            if m.ndim == 2 and m.shape[0] == nd:
                self.cart = m
                self.n_sensor = m.shape[1]
                self.sensor_idx = None
            elif m.size == 1 or m.size == numel:
                b = np.asarray(mask, dtype=bool).ravel()
                if b.size == 1:
                    b = np.full(numel, bool(b[0]))
                self.sensor_idx = np.where(b)[0]
                self.n_sensor = self.sensor_idx.size
            else:
                raise ValueError(f"Sensor mask shape {m.shape} is neither binary nor Cartesian")
        rec = getattr(self.sensor, "record", None) or ("p", "p_final")
        if isinstance(rec, str):
            rec = (rec,)
        rec = set(rec)
        ax = "xyz"[:nd]
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
        self.requested = set(rec)
        for k in list(rec):
            for suf in ("_max", "_min", "_rms"):
                if k.endswith(suf):
                    rec.add(k[: -len(suf)])
        if any(k.startswith("I") for k in self.requested):
            rec.update(["p"] + [f"u{a}" for a in ax])
        self.record = rec
        raw = getattr(self.sensor, "record_start_index", None)
        start = int(raw) if raw is not None else 1
        if start < 1 or start > self.Nt:
            raise ValueError("record_start_index out of range")
        self.start = start - 1
        self.n_rec = self.Nt - self.start

    def _extract(self, f):
        """Binary index gather or linear interpolation; kspace_solver.py:274, 328-361."""
        if self.cart is None:
            return f.ravel()[self.sensor_idx]
        axes = [(np.arange(N) - N // 2) * d for N, d in zip(self.shape, self.spacing)]
        if self.ndim == 1:
            return np.interp(self.cart.ravel(), axes[0], f.ravel())
        from scipy.interpolate import RegularGridInterpolator

        rgi = RegularGridInterpolator(axes, f, method="linear", bounds_error=True)
        return rgi(self.cart.T)

    def _setup_pml(self):
        """kspace_solver.py:363-406."""
        self.pml, self.pml_sg, self.pml_sizes = [], [], []
        for ax in range(self.ndim):
            N, dx = self.shape[ax], self.spacing[ax]
            name = "xyz"[ax]
            size = int(self._pml_size[ax]) if self._pml_size is not None else int(getattr(self.kgrid, f"pml_size_{name}", 0))
            alpha = float(self._pml_alpha[ax]) if self._pml_alpha is not None else float(getattr(self.kgrid, f"pml_alpha_{name}", 0))
            self.pml_sizes.append(size if alpha != 0 else 0)
            bshape = [1] * self.ndim
            bshape[ax] = N
            if size == 0 or alpha == 0:
                a = np.ones(N)
                b = np.ones(N)
            else:
                a = get_pml(N, dx, self.dt, self.c_ref, size, alpha, False)
                b = get_pml(N, dx, self.dt, self.c_ref, size, alpha, True)
            self.pml.append(a.reshape(bshape))
            self.pml_sg.append(b.reshape(bshape))

    def _setup_kspace(self):
        """kspace_solver.py:408-447."""
        self.k = []
        for ax, (N, dx) in enumerate(zip(self.shape, self.spacing)):
            bshape = [1] * self.ndim
            bshape[ax] = N
            self.k.append((2 * np.pi * np.fft.fftfreq(N, d=dx)).reshape(bshape))
        kmag = np.sqrt(sum(k ** 2 for k in self.k))
        self.kmag = kmag
        if self.use_kspace:
            arg = self.c_ref * kmag * self.dt / 2
            self.kappa = np.sinc(arg / np.pi)
            self.source_kappa = np.cos(arg)
        else:
            self.kappa = 1.0
            self.source_kappa = 1.0
        self.op_grad, self.op_div = [], []
        for ax, dx in enumerate(self.spacing):
            k = self.k[ax]
            if self.use_sg:
                self.op_grad.append(1j * k * np.exp(1j * k * dx / 2) * self.kappa)
                self.op_div.append(1j * k * np.exp(-1j * k * dx / 2) * self.kappa)
            else:
                self.op_grad.append(1j * k * self.kappa)
                self.op_div.append(1j * k * self.kappa)

    def _frac_lap(self, power):
        """kspace_solver.py:811-816."""
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.where(self.kmag == 0, 0, self.kmag ** power)

    def _setup_physics(self):
        """kspace_solver.py:449-550."""
        m = self.medium
        self.tau = self.eta = self.nabla1 = self.nabla2 = None
        self.stokes = False
        mode = getattr(m, "alpha_mode", None)
        mode = getattr(mode, "value", mode)
        if _is_enabled(getattr(m, "alpha_coeff", 0)):
            y = float(np.array(getattr(m, "alpha_power", 1.5) if getattr(m, "alpha_power", None) is not None else 1.5).flatten()[0])
            a_np = db2neper(_expand(m.alpha_coeff, self.shape, "alpha_coeff"), y)
            self.alpha_power = y
            if mode != "no_absorption":
                if abs(y - 2.0) < 1e-10:
                    self.stokes = True
                    self.tau = -2 * a_np * self.c0
                else:
                    self.tau = -2 * a_np * self.c0 ** (y - 1)
                    self.nabla1 = self._frac_lap(y - 2)
            if mode not in ("no_absorption", "no_dispersion") and abs(y - 2.0) >= 1e-10:
                self.eta = 2 * a_np * self.c0 ** y * np.tan(np.pi * y / 2)
                self.nabla2 = self._frac_lap(y - 1)
        bona = getattr(m, "BonA", 0)
        self.BonA = _expand(bona, self.shape, "BonA") if _is_enabled(bona) else None

    def _source_op(self, mask_raw, sig_raw, mode, scale):
        """kspace_solver.py:60-114."""
        mask = np.array(mask_raw, dtype=bool).ravel()
        if mask.size == 1:
            mask = np.full(int(np.prod(self.shape)), bool(mask[0]))
        n_src = int(mask.sum())
        s = np.array(sig_raw, dtype=np.float64)
        if s.ndim == 1:
            s = s.reshape(1, -1)
        elif s.ndim > 2:
            s = s.reshape(-1, s.shape[-1])
        scaled = s * np.atleast_1d(np.asarray(scale, dtype=np.float64))[:, None]
        L = scaled.shape[1]

        def val(t):
            if scaled.shape[0] == 1:
                return np.full(n_src, scaled[0, t])
            return scaled[:, t]

        if mode == "dirichlet":
            def op(t, f):
                if t >= L:
                    return f
                fl = f.flatten()
                fl[mask] = val(t)
                return fl.reshape(self.shape)
        elif mode == "additive":
            def op(t, f):
                if t >= L:
                    return f
                buf = np.zeros(mask.size)
                buf[mask] = val(t)
                return f + self._diff(buf.reshape(self.shape), self.source_kappa)
        elif mode == "additive-no-correction":
            def op(t, f):
                if t >= L:
                    return f
                buf = np.zeros(mask.size)
                buf[mask] = val(t)
                return f + buf.reshape(self.shape)
        else:
            raise ValueError(f"Unknown source mode: {mode!r}")
        return op

    def _setup_sources(self):
        """kspace_solver.py:552-645."""
        nd = self.ndim
        noop = lambda t, f: f  # noqa: E731
        src = self.source

        def c0_at(mask_raw):
            mask = np.array(mask_raw, dtype=bool).ravel()
            if mask.size == 1:
                mask = np.full(self.c0.size, bool(mask[0]))
            return self.c0.ravel()[mask]

        p_mask, p_sig = getattr(src, "p_mask", 0), getattr(src, "p", 0)
        p_mode = getattr(src, "p_mode", None) or "additive"
        if _is_enabled(p_mask) and _is_enabled(p_sig):
            c = c0_at(p_mask)
            if p_mode == "dirichlet":
                op = self._source_op(p_mask, p_sig, p_mode, 1.0 / (nd * c ** 2))
                self.src_p = [op] * nd
            else:
                self.src_p = [self._source_op(p_mask, p_sig, p_mode, 2 * self.dt / (nd * c * self.spacing[i])) for i in range(nd)]
        else:
            self.src_p = [noop] * nd
        u_mask = getattr(src, "u_mask", 0)
        u_mode = getattr(src, "u_mode", None) or "additive"
        self.src_u = []
        for i, name in enumerate(["ux", "uy", "uz"][:nd]):
            sig = getattr(src, name, 0)
            if _is_enabled(u_mask) and _is_enabled(sig):
                c = c0_at(u_mask)
                scale = np.ones_like(c) if u_mode == "dirichlet" else 2 * c * self.dt / self.spacing[i]
                self.src_u.append(self._source_op(u_mask, sig, u_mode, scale))
            else:
                self.src_u.append(noop)

    @staticmethod
    def _stagger(arr, axis):
        """kspace_solver.py:798-809."""
        lo = [slice(None)] * arr.ndim
        hi = [slice(None)] * arr.ndim
        lo[axis], hi[axis] = slice(None, -1), slice(1, None)
        last = [slice(None)] * arr.ndim
        last[axis] = slice(-1, None)
        return np.concatenate([0.5 * (arr[tuple(lo)] + arr[tuple(hi)]), arr[tuple(last)]], axis=axis)

    def _setup_fields(self):
        """kspace_solver.py:647-689."""
        nd = self.ndim
        self.p = np.zeros(self.shape)
        self.u = [np.zeros(self.shape) for _ in range(nd)]
        self.rho = [np.zeros(self.shape) for _ in range(nd)]
        self.rho0_sg = [self._stagger(self.rho0, a) for a in range(nd)] if self.use_sg else [self.rho0] * nd
        self.c0_sq = self.c0 ** 2
        self.dt_rho0 = [self.dt / r for r in self.rho0_sg]
        self.data = {}
        if "p" in self.record:
            self.data["p"] = np.zeros((self.n_sensor, self.n_rec))
        for a in "xyz"[:nd]:
            for suf in ("", "_staggered"):
                if f"u{a}{suf}" in self.record:
                    self.data[f"u{a}{suf}"] = np.zeros((self.n_sensor, self.n_rec))
        self.unstagger = [np.exp(-1j * self.k[a] * self.spacing[a] / 2) for a in range(nd)]
        p0 = getattr(self.source, "p0", 0)
        if _is_enabled(p0):
            p0 = _expand(p0, self.shape, "p0")
            if self.smooth_p0 and nd >= 2:
                p0 = smooth(p0, restore_max=True)
            self.p0 = p0
        else:
            self.p0 = None

    # -- step: kspace_solver.py:691-747 ---------------------------------------
    def _diff(self, f, op):
        """kspace_solver.py:781-796."""
        return np.real(self._ifftn(op * self._fftn(f)))

    def _mass_factor(self):
        """(2 sum(rho_i) + rho0) / rho0 of the nonlinear mass equation, 1 when linear;
        kspace_solver.py:549-550, 717."""
        return (2 * sum(self.rho) + self.rho0) / self.rho0 if self.BonA is not None else 1.0

    def step(self):
        if self.t >= self.Nt:
            return self
        nd = self.ndim
        P = self._fftn(self.p)
        for i in range(nd):
            g = np.real(self._ifftn(self.op_grad[i] * P))
            self.u[i] = self.pml_sg[i] * (self.pml_sg[i] * self.u[i] - self.dt_rho0[i] * g)
            self.u[i] = self.src_u[i](self.t, self.u[i])
        nl = self._mass_factor()
        div_total = np.zeros(self.shape)
        for i in range(nd):
            d = self._diff(self.u[i], self.op_div[i])
            div_total += d
            self.rho[i] = self.pml[i] * (self.pml[i] * self.rho[i] - self.dt * self.rho0 * d * nl)
            self.rho[i] = self.src_p[i](self.t, self.rho[i])
        rho_t = sum(self.rho)
        absorb = 0.0
        if self.tau is not None:
            absorb = self.tau * self.rho0 * div_total if self.stokes else self.tau * self._diff(self.rho0 * div_total, self.nabla1)
        disp = self.eta * self._diff(rho_t, self.nabla2) if self.eta is not None else 0.0
        nonlin = self.BonA * rho_t ** 2 / (2 * self.rho0) if self.BonA is not None else 0.0
        self.p = self.c0_sq * (rho_t + absorb - disp + nonlin)
        if self.t == 0 and self.p0 is not None:
            self.p = self.p0.copy()
            for i in range(nd):
                self.rho[i] = self.p0 / (self.c0_sq * nd)
                self.u[i] = (self.dt_rho0[i] / 2) * self._diff(self.p, self.op_grad[i])
        if self.t >= self.start:
            j = self.t - self.start
            if "p" in self.data:
                self.data["p"][:, j] = self._extract(self.p)
            for i, a in enumerate("xyz"[:nd]):
                if f"u{a}" in self.data:
                    self.data[f"u{a}"][:, j] = self._extract(self._diff(self.u[i], self.unstagger[i]))
                if f"u{a}_staggered" in self.data:
                    self.data[f"u{a}_staggered"][:, j] = self._extract(self.u[i])
        self.t += 1
        return self

    def run(self):
        """kspace_solver.py:749-778."""
        while self.t < self.Nt:
            self.step()
        res = dict(self.data)
        res.update(compute_aggregates(res, self.ndim, self.record))
        if "p" in res and any(f"u{a}" in res for a in "xyz") and any(k.startswith("I") for k in self.requested):
            res.update(acoustic_intensity(res))
        interior = tuple(slice(s, N - s if s else None) for s, N in zip(self.pml_sizes, self.shape))
        if "p_final" in self.requested:
            res["p_final"] = self.p[interior].copy()
        for i, a in enumerate("xyz"[: self.ndim]):
            if f"u{a}_final" in self.requested:
                res[f"u{a}_final"] = self.u[i][interior].copy()
        return {k: v for k, v in res.items() if k in self.requested}


def compute_aggregates(result, ndim, record):
    """kspace_solver.py:824-837."""
    out = {}
    for pre in ["p"] + [f"u{a}" for a in "xyz"[:ndim]]:
        ts = result.get(pre)
        if ts is None:
            continue
        if f"{pre}_max" in record:
            out[f"{pre}_max"] = np.max(ts, axis=-1)
        if f"{pre}_min" in record:
            out[f"{pre}_min"] = np.min(ts, axis=-1)
        if f"{pre}_rms" in record:
            out[f"{pre}_rms"] = np.sqrt(np.mean(ts ** 2, axis=-1))
    return out


def _half_step_shift(n):
    """Spectral multiplier advancing a length-n series by half a sample; kspace_solver.py:846-849."""
    freq = 2 * np.pi * np.arange(-(n // 2), n - n // 2) / n
    return np.fft.ifftshift(np.exp(1j * freq * 0.5))


def acoustic_intensity(result):
    """Temporal half-step Fourier shift of u, then I = p u; kspace_solver.py:840-862."""
    p = result["p"]
    shift = _half_step_shift(p.shape[-1])
    out = {}
    for a in "xyz":
        u = result.get(f"u{a}")
        if u is None:
            continue
        us = np.real(np.fft.ifft(shift * np.fft.fft(u, axis=-1), axis=-1))
        out[f"I{a}"] = p * us
        out[f"I{a}_avg"] = np.mean(out[f"I{a}"], axis=-1)
    return out
