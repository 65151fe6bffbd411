"""Owning wrapper of a device plan (a batch of same-shape solver instances)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _abi
from .._abi import check, f32, fptr


_ABSORBING = _abi.KW_STOKES | _abi.KW_POWERLAW_ABS | _abi.KW_POWERLAW_DISP


def _i64(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(C.POINTER(C.c_int64))


class DevicePlan:
    """A batch of instances that share grid shape, spacing and physics flags."""

    def __init__(self, setups, device=0):
        s0 = setups[0]
        for s in setups[1:]:
            if s.shape != s0.shape or list(s.spacing) != list(s0.spacing):
                raise ValueError("all instances of a plan need the same grid shape and spacing")
            if s.flags != s0.flags or s.use_sg != s0.use_sg:
                raise ValueError("all instances of a plan need the same physics switches")
            if (s.p0 is None) != (s0.p0 is None):
                raise ValueError("either every instance of a plan has p0 or none has")
            if (s0.flags & _ABSORBING) and s.alpha_power != s0.alpha_power:
                # the fractional-Laplacian exponents y-2, y-1 and tan(pi y / 2) are set once per plan
                raise ValueError("all absorbing instances of a plan need the same alpha_power "
                                 f"(got {s0.alpha_power} and {s.alpha_power})")
        L = _abi.lib()
        d = _abi.PlanDesc()
        d.ndim = s0.ndim
        for i in range(3):
            d.n[i] = s0.shape[i] if i < s0.ndim else 1
            d.spacing[i] = s0.spacing[i] if i < s0.ndim else 1.0
        d.batch = len(setups)
        d.flags = s0.flags
        d.alpha_power = s0.alpha_power
        d.use_sg = int(s0.use_sg)
        d.device = int(device)
        self._h = C.c_void_p()
        check(L.kw_plan_create(C.byref(d), C.byref(self._h)))
        self.setups = setups
        self.ndim = s0.ndim
        self.shape = s0.shape
        self._rec_flags = []
        try:
            for b, s in enumerate(setups):
                self._upload(b, s)
            check(L.kw_plan_init(self._h))
        except Exception:
            self.close()
            raise

    # -- upload -------------------------------------------------------------
    def _field(self, b, which, val):
        arr = f32(np.atleast_1d(val)).ravel()
        check(_abi.lib().kw_plan_set_field(self._h, b, which, fptr(arr), arr.size))

    def _upload(self, b, s):
        L = _abi.lib()
        inst = _abi.InstDesc()
        inst.nt, inst.dt, inst.c_ref, inst.use_kspace = s.Nt, s.dt, s.c_ref, int(s.use_kspace)
        keep = []
        for ax in range(s.ndim):
            a, g = f32(s.pml[ax]), f32(s.pml_sg[ax])
            keep += [a, g]
            inst.pml[ax] = fptr(a)
            inst.pml_sg[ax] = fptr(g)
        check(L.kw_plan_set_instance(self._h, b, C.byref(inst)))
        if s.p0_split is not None:
            share = (C.c_double * 3)(*(tuple(s.p0_split) + (0.0,) * 3)[:3])
            check(L.kw_plan_set_density_split(self._h, b, share))
        self._field(b, _abi.F_C0SQ, np.asarray(s.c0) ** 2)
        self._field(b, _abi.F_RHO0, s.rho0)
        for ax in range(s.ndim):
            self._field(b, _abi.F_DTR0_X + ax, s.dtr0[ax])
        if s.bona is not None:
            self._field(b, _abi.F_BONA, s.bona)
        if s.alpha_np is not None and s.flags & (_abi.KW_STOKES | _abi.KW_POWERLAW_ABS | _abi.KW_POWERLAW_DISP):
            self._field(b, _abi.F_ALPHA_NP, s.alpha_np)
        if s.tau is not None:
            self._field(b, _abi.F_TAU, s.tau)
        if s.eta is not None:
            self._field(b, _abi.F_ETA, s.eta)
        if s.p0 is not None:
            self._field(b, _abi.F_P0, s.p0)
        for which, term in [(-1, s.src_p)] + list(enumerate(s.src_u)):
            if term is None:
                continue
            idx, ip = _i64(term.idx)
            sc = f32(term.scale)
            sr = np.ascontiguousarray(term.sigrow, dtype=np.int32)
            sig = f32(term.sig)
            fac = (C.c_double * 3)(*term.fac[:3])
            check(L.kw_plan_set_source(self._h, b, which, term.mode, idx.size, ip, fptr(sc),
                                       sr.ctypes.data_as(C.POINTER(C.c_int32)), fptr(sig), sig.shape[0], sig.shape[1], fac))
        rec = s.record
        flags = 0
        want_series = "p" in rec
        if want_series:
            flags |= _abi.REC_P
        for i, a in enumerate("xyz"[: s.ndim]):
            if f"u{a}_staggered" in rec:
                flags |= _abi.REC_UX << i
            if f"u{a}" in rec:
                flags |= _abi.REC_UCX << i
        self._rec_flags.append(flags)
        if flags:
            sidx, sp = _i64(s.sensor_idx)
            w = f32(s.sensor_w) if s.sensor_w is not None else None
            check(L.kw_plan_set_sensor(self._h, b, s.n_sensor, sidx.shape[1], sp, fptr(w) if w is not None else None,
                                       s.rec_start, flags))

    # -- run / read ---------------------------------------------------------
    def run(self, nsteps):
        check(_abi.lib().kw_plan_run(self._h, int(nsteps)))

    def launch(self, nsteps):
        """Enqueue nsteps steps without waiting (see sync())."""
        check(_abi.lib().kw_plan_launch(self._h, int(nsteps)))

    def sync(self):
        check(_abi.lib().kw_plan_sync(self._h))

    def run_profiled(self, nsteps):
        """Run nsteps with per-kernel HIP-event timing; returns {kind: (ms, bytes, launches)}."""
        L = _abi.lib()
        nk = L.kw_kernel_kind_count()
        ms = np.zeros(nk)
        by = np.zeros(nk)
        nl = np.zeros(nk, dtype=np.int64)
        check(L.kw_plan_run_profiled(self._h, int(nsteps), ms.ctypes.data_as(C.POINTER(C.c_double)),
                                     by.ctypes.data_as(C.POINTER(C.c_double)), nl.ctypes.data_as(C.POINTER(C.c_int64))))
        return {L.kw_kernel_kind_name(k).decode(): (float(ms[k]), float(by[k]), int(nl[k])) for k in range(nk) if nl[k]}

    def reset(self):
        """Zero the state and re-apply the t = 0 initial conditions."""
        check(_abi.lib().kw_plan_init(self._h))

    def read(self, b, which):
        n = C.c_int64(0)
        L = _abi.lib()
        check(L.kw_plan_field_size(self._h, b, which, C.byref(n)))
        out = np.empty(n.value, dtype=np.float32)
        check(L.kw_plan_read(self._h, b, which, fptr(out), out.size))
        return out

    def tiles(self):
        """(FFT-only, spectral-job) column tile variants the plan picked by timing (0 default, 1 half width)."""
        f, j = C.c_int(0), C.c_int(0)
        check(_abi.lib().kw_plan_tiles(self._h, C.byref(f), C.byref(j)))
        return f.value, j.value

    def radix_plans(self):
        """(FFT-only, spectral-job) radix-plan variants the plan picked by timing (0 default, 1 radix 4
        next to a factor 5 or 7)."""
        f, j = C.c_int(0), C.c_int(0)
        check(_abi.lib().kw_plan_radix_plans(self._h, C.byref(f), C.byref(j)))
        return f.value, j.value

    def step_index(self, b=0):
        return int(self.read(b, _abi.F_STEP)[0])

    def recorded(self, b, which):
        """Recorded series as (n_sensor, n_recorded_steps)."""
        s = self.setups[b]
        return self.read(b, which).reshape(s.Nt - s.rec_start, s.n_sensor).T

    def field(self, b, which):
        return self.read(b, which).reshape(self.shape)

    def adjoint_c(self, adj, checkpoint=50):
        """dJ/dc0 per instance by the reverse-time adjoint (kw_plan_adjoint_c).

        adj[b] is dJ/d(recorded p) of instance b, shaped like recorded(b, F_REC_P)
        (n_sensor, n_recorded_steps).  Returns a list of grid-shaped float64
        gradients.  Re-runs the forward problem from t = 0; call reset() before
        stepping the plan again.
        """
        rows = []
        for b, s in enumerate(self.setups):
            a = np.asarray(adj[b], dtype=np.float32)
            if a.shape != (s.n_sensor, s.Nt - s.rec_start):
                raise ValueError(f"adjoint source of instance {b} has shape {a.shape}, expected {(s.n_sensor, s.Nt - s.rec_start)}")
            rows.append(a.T)
        a = np.ascontiguousarray(np.stack(rows), dtype=np.float32)
        out = np.empty((len(self.setups),) + tuple(self.shape), dtype=np.float32)
        check(_abi.lib().kw_plan_adjoint_c(self._h, fptr(a), int(checkpoint), fptr(out)))
        return [out[b].astype(np.float64) for b in range(len(self.setups))]

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _abi.lib().kw_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
