"""Drop-in ``Simulation`` for the native solver path, running on the MI355X.

Mirrors ``kwave.solvers.kspace_solver.Simulation`` (kwave/solvers/
kspace_solver.py:122-778): same constructor, ``setup()``, ``step()``,
``run()`` and result dictionary.  Set-up is done on the host in float64
(solvers/setup.py); every time step runs in the HIP kernels of libkwgpu.so.
Arithmetic in the loop is float32; results are returned in the requested
dtype (float64 by default, like the reference).
"""
from __future__ import annotations

import os
import warnings

import numpy as np

from .. import _abi
from . import embed1d
from .plan import DevicePlan
from .setup import build_instance


def _resolve_dtype(dtype):
    """kspace_solver.py:161-179."""
    if dtype is None:
        return np.float64
    if dtype in (np.float32, np.float64):
        return dtype
    try:
        resolved = np.dtype(dtype).type
    except TypeError as e:
        raise ValueError(f"dtype must be np.float32 or np.float64 (or string equivalent), got {dtype!r}") from e
    if resolved not in (np.float32, np.float64):
        raise ValueError(f"dtype must resolve to float32 or float64, got {resolved.__name__}")
    return resolved


def compute_aggregates(result, ndim, record):
    """max/min/rms over the recorded window (kspace_solver.py:824-837)."""
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


def acoustic_intensity(result):
    """Intensity from the recorded series (kspace_solver.py:840-862): the velocity is
    shifted forward by dt/2 with a temporal Fourier interpolant, then I = p u."""
    p = result["p"]
    n = p.shape[-1]
    freq = 2 * np.pi * np.arange(-(n // 2), n - n // 2) / n
    shift = np.fft.ifftshift(np.exp(1j * freq * 0.5))
    out = {}
    for a in "xyz":
        u = result.get(f"u{a}")
        if u is None:
            continue
        us = np.real(np.fft.ifft(shift * np.fft.fft(u, axis=-1), axis=-1))
        out[f"I{a}"] = p * us
        out[f"I{a}_avg"] = np.mean(out[f"I{a}"], axis=-1)
    return out


def collect_results(plan, b, st, dtype):
    """Assemble the reference's result dict for instance b (kspace_solver.py:762-778)."""
    result = {}
    if "p" in st.record:
        result["p"] = plan.recorded(b, _abi.F_REC_P)
    for i, a in enumerate("xyz"[: st.ndim]):
        if f"u{a}" in st.record:
            result[f"u{a}"] = plan.recorded(b, _abi.F_REC_UCX + i)
        if f"u{a}_staggered" in st.record:
            result[f"u{a}_staggered"] = plan.recorded(b, _abi.F_REC_UX + i)
    result.update(compute_aggregates(result, st.ndim, st.record))
    if "p" in result and any(f"u{a}" in result for a in "xyz") and any(k.startswith("I") for k in st.requested):
        result.update(acoustic_intensity(result))
    interior = tuple(slice(s, N - s if s else None) for s, N in zip(st.pml_sizes, st.shape))
    if "p_final" in st.requested:
        result["p_final"] = plan.field(b, _abi.F_P)[interior].copy()
    for i, a in enumerate("xyz"[: st.ndim]):
        if f"u{a}_final" in st.requested:
            result[f"u{a}_final"] = plan.field(b, _abi.F_UX + i)[interior].copy()
    return {k: np.asarray(v, dtype=dtype) for k, v in result.items() if k in st.requested}


class Simulation:
    """N-D k-space pseudospectral propagator with setup/step separation.

    ``device`` selects the HIP device ("gpu", "gpu:N" or an int); "cpu" is
    accepted for call compatibility with the reference and also runs on the
    GPU, since this package has no CPU solver.
    """

    def __init__(self, kgrid, medium, source, sensor, *, device="gpu", use_sg=True, use_kspace=True, smooth_p0=True,
                 pml_size=None, pml_alpha=None, quiet=False, dtype=None):
        self.kgrid, self.medium, self.source, self.sensor = kgrid, medium, source, sensor
        self.use_sg, self.use_kspace, self.smooth_p0, self.quiet = use_sg, use_kspace, smooth_p0, quiet
        self._pml_size_override = pml_size
        self._pml_alpha_override = pml_alpha
        self._dtype = _resolve_dtype(dtype)
        if pml_size is None and hasattr(kgrid, "N") and hasattr(kgrid, "spacing"):
            warnings.warn(
                "Simulation received a kWaveGrid without explicit pml_size; PML will be disabled. Pass "
                "pml_size=(20,)*kgrid.dim or use kspaceFirstOrder() which handles PML automatically.",
                UserWarning, stacklevel=2)
        self._device = self._device_index(device)
        self._embed1d = False
        self._is_setup = False
        self._plan = None
        self._t = 0

    @staticmethod
    def _device_index(device):
        if isinstance(device, int):
            return device
        if device in ("gpu", "cpu", None):
            return int(os.environ.get("KW_DEVICE", "0"))
        if isinstance(device, str) and device.startswith("gpu:"):
            return int(device.split(":", 1)[1])
        raise ValueError(f"Unknown device {device!r}. Use 'gpu' or 'cpu'.")

    # -- lifecycle ------------------------------------------------------------
    def setup(self):
        """Build the host tables and the device plan (kspace_solver.py:205-241).  A 1-D grid
        runs as a 2 x N grid uniform along axis 0 (solvers/embed1d.py)."""
        self._embed1d = embed1d.is_1d(self.kgrid)
        if self._embed1d:
            kg, med, src, sen, pml, alpha = embed1d.embed(self.kgrid, self.medium, self.source, self.sensor,
                                                          self._pml_size_override, self._pml_alpha_override)
            # the reference does not smooth p0 in 1-D (kspace_solver.py:682)
            self._setup = embed1d.adjust_setup(build_instance(kg, med, src, sen, use_sg=self.use_sg,
                                                              use_kspace=self.use_kspace, smooth_p0=False,
                                                              pml_size=pml, pml_alpha=alpha))
        else:
            self._setup = build_instance(self.kgrid, self.medium, self.source, self.sensor, use_sg=self.use_sg,
                                         use_kspace=self.use_kspace, smooth_p0=self.smooth_p0,
                                         pml_size=self._pml_size_override, pml_alpha=self._pml_alpha_override)
        st = self._setup
        self.grid_shape, self.spacing, self.ndim = st.shape, st.spacing, st.ndim
        if self._embed1d:
            self.grid_shape, self.spacing, self.ndim = st.shape[1:], st.spacing[1:], 1
        self.Nt, self.dt, self.c_ref = st.Nt, st.dt, st.c_ref
        self.pml_sizes = st.pml_sizes[1:] if self._embed1d else st.pml_sizes
        self.record = st.record
        self.n_sensor_points = st.n_sensor
        self.record_start_index = st.rec_start
        self.num_recorded_time_points = st.Nt - st.rec_start
        self._plan = DevicePlan([st], device=self._device)
        # the device applies the t = 0 initial-value step during plan set-up;
        # the host counter starts at 0 like the reference and catches up in step(),
        # and the state properties show the reference's zeros until then
        self._t = 0
        self._is_setup = True
        return self

    @property
    def t(self):
        return self._t

    def step(self):
        """Advance one time step on the device (kspace_solver.py:691-747)."""
        if not self._is_setup:
            self.setup()
        if self._t >= self.Nt:
            return self
        if self._plan.step_index(0) <= self._t:
            self._plan.run(1)
        self._t += 1
        return self

    def run(self):
        """Run to completion and return the result dict (kspace_solver.py:749-778)."""
        if not self._is_setup:
            self.setup()
        remaining = self.Nt - self._plan.step_index(0)
        if remaining > 0:
            self._plan.run(remaining)
        self._t = self._plan.step_index(0)
        res = collect_results(self._plan, 0, self._setup, self._dtype)
        return embed1d.unembed(res) if self._embed1d else res

    # -- state inspection (reads device memory) ------------------------------
    # Before the first step() the reference's fields are the zeros of
    # _setup_fields (kspace_solver.py:651-654); the device plan has already
    # applied the t = 0 initial-value override by then, so report zeros.
    def _state(self, which):
        if self._t == 0:
            return np.zeros(self.grid_shape, dtype=self._dtype)
        f = self._plan.field(0, which)
        return (f[embed1d.ROW] if self._embed1d else f).astype(self._dtype)

    @property
    def p(self):
        return self._state(_abi.F_P)

    @property
    def u(self):
        if self._embed1d:
            return [self._state(_abi.F_UX + 1)]  # the propagation axis of the embedding
        return [self._state(_abi.F_UX + i) for i in range(self.ndim)]

    @property
    def rho_split(self):
        if self._embed1d:  # the 1-D density is the sum of the embedding's split densities
            return [(self._state(_abi.F_RHOX) + self._state(_abi.F_RHOX + 1)).astype(self._dtype)]
        return [self._state(_abi.F_RHOX + i) for i in range(self.ndim)]

    def close(self):
        if self._plan is not None:
            self._plan.close()
            self._plan = None
