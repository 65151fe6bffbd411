"""The k-Wave C++ binaries' file route, run on the MI355X.

Mirrors ``kwave.solvers.cpp_simulation.CppSimulation`` (kwave/solvers/
cpp_simulation.py:33-136): ``prepare()`` writes the binaries' HDF5 input file,
``run()`` executes it and returns the result dict in the same C order.  The
binary (kspaceFirstOrder-OMP / -CUDA) is replaced by this package's runner
(``kwave_amd.solvers.cpp_runner``, also a command line with the binaries'
-i / -o options), which reads the input file, runs the time loop on the GPU and
writes the output file; a file written by another tool in the same layout runs
the same way.

Input variables (``input_variables``) follow cpp_simulation.py:139-312: grid
sizes and spacings, Nt and dt, c0 / rho0 / rho0_sg{x,y,z} (scalars or
grids), c_ref = max(c0), PML sizes and alphas, absorption (alpha_coeff in
dB/(MHz^y cm), alpha_power, absorbing_flag 0 / 1 power law / 2 Stokes),
BonA, the source and sensor flags, 1-based Fortran-order index lists
(sensor_mask_index, p_source_index, u_source_index) and the source signals.
Two choices differ from that writer and follow the binaries' input semantics
instead, as the reference's legacy writer (kWaveSimulation_helper/
save_to_disk_func.py with scale_source_terms_func.py:89-191, pinned by
tests/reference_outputs/out_tvsp_homogeneous_medium_monopole.json) does:

* time-varying sources are stored scaled to the units the time loop adds:
  additive pressure x 2 dt / (D c0 dx), Dirichlet pressure x 1 / (D c0^2),
  additive velocity x 2 c0 dt / d_i (with the sound speed at each source
  point; a single signal row becomes one row per point when c0 varies there);
* rho0_sg{x,y,z} are the staggered-grid densities of the time loop (midpoint
  average keeping the last plane, kspace_solver.py:798-809) rather than a
  periodic roll.
"""
from __future__ import annotations

import os
import shutil
import tempfile
import warnings
from collections import OrderedDict

import numpy as np

from ..utils import h5io
from .setup import stagger

SOURCE_MODE = {"dirichlet": 0, "additive-no-correction": 1, "additive": 2}
SOURCE_MODE_NAME = {v: k for k, v in SOURCE_MODE.items()}
_FULL_GRID_SUFFIXES = ("_final", "_max_all", "_min_all", "_rms_all")


def _f_order_points(mask):
    """Flat C-order indices of the mask's points, listed in Fortran order (the file's order)."""
    shape = mask.shape
    f_flat = np.where(mask.ravel(order="F"))[0]
    return np.ravel_multi_index(np.unravel_index(f_flat, shape, order="F"), shape), f_flat


def _signal_rows(sig, npts):
    s = np.asarray(sig, dtype=np.float64)
    if s.ndim == 1:
        s = s.reshape(1, -1)
    elif s.ndim > 2:
        s = s.reshape(-1, s.shape[-1])
    if s.shape[0] not in (1, npts):
        raise ValueError(f"source signal has {s.shape[0]} rows for {npts} source points")
    return s


def _grid_or_scalar(val, shape):
    a = np.asarray(val, dtype=np.float64)
    return a.reshape(shape) if a.size > 1 else a.reshape(1)


def input_variables(kgrid, medium, source, sensor, pml_size, pml_alpha, use_sg=True):
    """The variables of the binaries' input file, as C-ordered arrays (float32 or uint64), in file order."""
    nd = int(kgrid.dim)
    shape = tuple(int(n) for n in np.atleast_1d(kgrid.N))
    spacing = [float(d) for d in np.atleast_1d(kgrid.spacing)]
    dt, nt = float(kgrid.dt), int(kgrid.Nt)
    c0 = _grid_or_scalar(medium.sound_speed, shape)
    rho = getattr(medium, "density", None)
    rho0 = _grid_or_scalar(1000.0 if rho is None else rho, shape)
    c_ref = float(np.max(c0))
    if rho0.size == 1 or not use_sg:
        rho0_sg = [rho0] * nd
    else:
        rho0_sg = [stagger(rho0, a) for a in range(nd)]
    numel = int(np.prod(shape))

    def mask_of(m, what):
        a = np.asarray(m)
        if a.ndim == 2 and a.shape[0] == nd and a.size != numel:
            raise ValueError(f"Cartesian {what} masks are not supported by the binaries' file format; use a binary "
                             "(grid-shaped) mask")
        a = a.astype(bool).reshape(-1)
        if a.size == 1:
            a = np.full(numel, bool(a[0]))
        return a.reshape(shape)

    u64 = lambda v: np.asarray(v, dtype=np.uint64)  # noqa: E731
    f32 = lambda v: np.asarray(v, dtype=np.float32)  # noqa: E731
    has_p0 = getattr(source, "p0", None) is not None and np.any(np.asarray(source.p0) != 0)
    has_p = getattr(source, "p", None) is not None and np.any(np.asarray(source.p) != 0)
    vel = [n for n in ("ux", "uy", "uz")[:nd] if getattr(source, n, None) is not None]
    if sensor is None or getattr(sensor, "mask", None) is None:
        smask = np.ones(shape, dtype=bool)
    else:
        smask = mask_of(sensor.mask, "sensor")
    bona = getattr(medium, "BonA", None)
    nonlinear = bona is not None and np.any(np.asarray(bona) != 0)
    alpha = getattr(medium, "alpha_coeff", None)
    y = getattr(medium, "alpha_power", None)
    if alpha is None or np.all(np.asarray(alpha) == 0):
        absorbing = 0
    elif y is not None and abs(float(np.asarray(y).flat[0]) - 2.0) < 1e-10:
        absorbing = 2
    else:
        absorbing = 1

    v = OrderedDict()
    v["Nx"], v["Ny"], v["Nz"], v["Nt"] = u64(shape[0]), u64(shape[1] if nd > 1 else 1), u64(shape[2] if nd > 2 else 1), u64(nt)
    for i, a in enumerate("xyz"):
        v[f"pml_{a}_size"] = u64(int(pml_size[i]) if i < nd else 0)
    # sources (values scaled to the units the time loop injects; see the module docstring)
    p_rows = u_rows = None
    if has_p:
        pm = mask_of(source.p_mask, "source")
        c_idx, f_idx = _f_order_points(pm)
        rows = _signal_rows(source.p, c_idx.size)
        mode = getattr(source, "p_mode", None) or "additive"
        c_at = c0.reshape(-1)[c_idx] if c0.size > 1 else np.full(c_idx.size, c0[0])
        if mode == "dirichlet":
            scale = 1.0 / (nd * c_at ** 2)
        else:
            scale = 2.0 * dt / (nd * c_at * spacing[0])
        if rows.shape[0] > 1:
            rows = rows[np.argsort(np.argsort(c_idx))]  # signal rows arrive in C order of the points
        if rows.shape[0] == 1 and np.ptp(scale) == 0:
            p_rows = rows * scale[0]
        else:
            p_rows = np.broadcast_to(rows, (c_idx.size, rows.shape[1])) * scale[:, None]
        p_mode, p_idx = mode, f_idx
    if vel:
        um = mask_of(source.u_mask, "source")
        c_idx, f_idx = _f_order_points(um)
        mode = getattr(source, "u_mode", None) or "additive"
        c_at = c0.reshape(-1)[c_idx] if c0.size > 1 else np.full(c_idx.size, c0[0])
        u_rows = {}
        for name in vel:
            rows = _signal_rows(getattr(source, name), c_idx.size)
            if rows.shape[0] > 1:
                rows = rows[np.argsort(np.argsort(c_idx))]
            if mode == "dirichlet":
                u_rows[name] = rows
            else:
                scale = 2.0 * c_at * dt / spacing["xyz".index(name[1])]
                u_rows[name] = rows * scale[0] if rows.shape[0] == 1 and np.ptp(scale) == 0 else \
                    np.broadcast_to(rows, (c_idx.size, rows.shape[1])) * scale[:, None]
        u_mode, u_idx = mode, f_idx
    v["p_source_flag"] = u64(p_rows.shape[1] if has_p else 0)
    v["p0_source_flag"] = u64(int(has_p0))
    for n in ("ux", "uy", "uz"):
        v[f"{n}_source_flag"] = u64(u_rows[n].shape[1] if vel and n in u_rows else 0)
    for n in ("sxx", "syy", "szz", "sxy", "sxz", "syz"):
        v[f"{n}_source_flag"] = u64(0)
    v["transducer_source_flag"] = u64(0)
    v["nonuniform_grid_flag"] = u64(0)
    v["nonlinear_flag"] = u64(int(nonlinear))
    v["absorbing_flag"] = u64(absorbing)
    v["elastic_flag"] = u64(0)
    v["axisymmetric_flag"] = u64(0)
    v["sensor_mask_type"] = u64(0)
    v["sensor_mask_index"] = u64(np.where(smask.ravel(order="F"))[0] + 1).reshape(-1, 1)
    v["dx"], v["dt"] = f32(spacing[0]), f32(dt)
    v["c0"], v["c_ref"], v["rho0"] = f32(c0 if c0.size > 1 else c0), f32(c_ref), f32(rho0)
    if nd >= 2:
        v["dy"], v["pml_y_alpha"] = f32(spacing[1]), f32(pml_alpha[1])
    if nd >= 3:
        v["dz"], v["pml_z_alpha"] = f32(spacing[2]), f32(pml_alpha[2])
    v["pml_x_alpha"] = f32(pml_alpha[0])
    for i, a in enumerate("xyz"[:nd]):
        v[f"rho0_sg{a}"] = f32(rho0_sg[i])
    if nonlinear:
        v["BonA"] = f32(_grid_or_scalar(bona, shape))
    if alpha is not None:
        v["alpha_coeff"] = f32(_grid_or_scalar(alpha, shape))
    if y is not None:
        v["alpha_power"] = f32(float(np.asarray(y).flat[0]))
    if has_p0:
        v["p0_source_input"] = f32(np.asarray(source.p0, dtype=np.float64).reshape(shape))
    if has_p:
        v["p_source_input"] = f32(p_rows)
        v["p_source_index"] = u64(p_idx + 1).reshape(-1, 1)
        v["p_source_mode"] = u64(SOURCE_MODE.get(p_mode, 2))
        v["p_source_many"] = u64(int(p_rows.shape[0] > 1))
    if vel:
        for n in vel:
            v[f"{n}_source_input"] = f32(u_rows[n])
        v["u_source_index"] = u64(u_idx + 1).reshape(-1, 1)
        v["u_source_mode"] = u64(SOURCE_MODE.get(u_mode, 2))
        v["u_source_many"] = u64(int(any(r.shape[0] > 1 for r in u_rows.values())))
    return v


def write_input_file(filename, variables):
    """Write the variables and the file attributes (kwave/utils/io.py:52-243)."""
    for name, val in variables.items():
        h5io.write_matrix(filename, val, name)
    h5io.write_attributes(filename, file_type="input")


class CppSimulation:
    """Serialise a simulation to the binaries' HDF5 input file and run it on the MI355X."""

    def __init__(self, kgrid, medium, source, sensor, *, pml_size, pml_alpha, use_sg=True):
        self.kgrid, self.medium, self.source, self.sensor = kgrid, medium, source, sensor
        self.pml_size, self.pml_alpha, self.use_sg = pml_size, pml_alpha, use_sg
        self.ndim = kgrid.dim

    def prepare(self, data_path=None):
        """Write the input file; returns (input_file, output_file)."""
        if data_path is None:
            data_path = tempfile.mkdtemp(prefix="kwave_")
        os.makedirs(data_path, exist_ok=True)
        input_file = os.path.join(data_path, "kwave_input.h5")
        output_file = os.path.join(data_path, "kwave_output.h5")
        if os.path.exists(input_file):
            raise FileExistsError(f"{input_file!r} already exists. Delete it or choose a different data_path to avoid "
                                  "overwriting previous simulation inputs.")
        write_input_file(input_file, input_variables(self.kgrid, self.medium, self.source, self.sensor, self.pml_size,
                                                     self.pml_alpha, self.use_sg))
        return input_file, output_file

    def run(self, *, device="gpu", num_threads=None, device_num=None, quiet=False, debug=False, data_path=None,
            binary_path=None, record=None):
        """prepare(), execute the file on the GPU (cpp_runner), read the output and return it in C order."""
        from .cpp_runner import run_input_file

        if binary_path is not None:
            warnings.warn("binary_path is ignored: the input file runs on the MI355X solver of this package", stacklevel=2)
        cleanup = data_path is None
        input_file, output_file = self.prepare(data_path=data_path)
        try:
            run_input_file(input_file, output_file, device=0 if device_num is None else int(device_num),
                           record=record or ("p",))
            return self._fix_output_order(self._parse_output(output_file))
        finally:
            if cleanup:
                shutil.rmtree(os.path.dirname(input_file), ignore_errors=True)

    @staticmethod
    def _parse_output(output_file):
        """Every variable of the output file, squeezed (kwave/executor.py:120-124)."""
        return {k: v.squeeze() for k, v in h5io.read_file(output_file).items()}

    def _fix_output_order(self, result):
        """Reversed-dimension grid fields -> C order; sensor rows from Fortran to C order of the mask points
        (kwave/solvers/cpp_simulation.py:93-135).  Per-sensor aggregates over time (p_max, p_min, p_rms:
        one value per sensor point) are reordered like the time series."""
        nd = self.ndim
        shape = tuple(int(n) for n in np.atleast_1d(self.kgrid.N))
        out = dict(result)
        for k, val in out.items():
            if isinstance(val, np.ndarray) and val.ndim == nd and any(k.endswith(s) for s in _FULL_GRID_SUFFIXES):
                out[k] = val.transpose(tuple(range(nd - 1, -1, -1)))
        if self.sensor is None or getattr(self.sensor, "mask", None) is None:
            mask = np.ones(shape, dtype=bool)
        else:
            mask = np.asarray(self.sensor.mask, dtype=bool).reshape(shape)
        n = int(mask.sum())
        if n > 0 and nd >= 2:
            c_idx, _ = _f_order_points(mask)
            perm = np.argsort(c_idx, kind="stable")  # file row j holds the j-th point in F order
            for k, val in out.items():
                if not isinstance(val, np.ndarray) or any(k.endswith(s) for s in _FULL_GRID_SUFFIXES):
                    continue
                if val.ndim in (1, 2) and val.shape[0] == n:
                    out[k] = val[perm]
        return out
