"""Run a k-Wave C++-binary input file on the MI355X and write the binary's output file.

The drop-in for ``kspaceFirstOrder-OMP`` / ``kspaceFirstOrder-CUDA`` on the file
route (kwave/solvers/cpp_simulation.py:412-429 builds their command line):

    python -m kwave_amd.solvers.cpp_runner -i input.h5 -o output.h5 [-g DEVICE]
        [--p_raw] [--p_max] [--p_min] [--p_rms] [--p_max_all] [--p_min_all] [--p_final]
        [--u_raw] [--u_non_staggered_raw] [--u_final]

With no output flag the pressure time series is recorded (--p_raw).  The input
variables are read as the binaries define them (see cpp_simulation.py in this
package): 1-based Fortran-order index lists, source signals already scaled to
the units the time loop injects, staggered densities as given.  The output file
holds, in the binaries' layout: time series with one row per sensor point in
Fortran order of the mask ('p', 'ux'/'uy'/'uz' for the staggered velocity,
'ux_non_staggered'... on the pressure grid), per-point aggregates ('p_max',
'p_min', 'p_rms'), and whole-grid fields ('p_final', 'ux_final'..., 'p_max_all',
'p_min_all') in reversed-dimension storage.  File type "output".
"""
from __future__ import annotations

import argparse
import sys

import numpy as np

from ..utils import h5io

RECORD_FLAGS = ("p_raw", "p_max", "p_min", "p_rms", "p_max_all", "p_min_all", "p_final", "u_raw", "u_non_staggered_raw",
                "u_final")


def _scalar(v, k, default=None):
    if k not in v:
        return default
    return np.asarray(v[k]).reshape(-1)[0]


def _field(v, k, shape, nd):
    a = np.asarray(v[k], dtype=np.float64)
    if a.size == 1:
        return float(a.reshape(-1)[0])
    return h5io.from_kwave_layout(a, nd).reshape(shape)


def _signals(v, k, nd):
    """(rows, length) signal matrix of a source variable (stored from a 2-D (rows, length) array)."""
    a = np.asarray(v[k], dtype=np.float64)
    return h5io.from_kwave_layout(a, 2) if a.size > 1 else a.reshape(1, 1)


def _mask_from_index(idx, shape):
    """Boolean C-ordered mask from 1-based Fortran-order linear indices."""
    f = np.asarray(idx, dtype=np.int64).reshape(-1) - 1
    m = np.zeros(int(np.prod(shape)), dtype=bool)
    m[np.ravel_multi_index(np.unravel_index(f, shape, order="F"), shape)] = True
    return m.reshape(shape)


def _c_rows(rows, idx, shape):
    """Signal rows listed in Fortran order of the points -> C order of the points."""
    if rows.shape[0] == 1:
        return rows
    f = np.asarray(idx, dtype=np.int64).reshape(-1) - 1
    c_idx = np.ravel_multi_index(np.unravel_index(f, shape, order="F"), shape)
    return rows[np.argsort(c_idx, kind="stable")]


def problem_from_file(v):
    """(kgrid, medium, source, sensor, pml_size, pml_alpha, rho0_sg) from the variables of an input file."""
    from ..kgrid import kWaveGrid
    from ..kmedium import kWaveMedium
    from ..ksensor import kSensor
    from ..ksource import kSource

    n = [int(_scalar(v, k, 1)) for k in ("Nx", "Ny", "Nz")]
    nd = 3 if n[2] > 1 else 2
    if n[1] <= 1:
        raise ValueError("1-D input files are not run by this package (2-D and 3-D grids)")
    shape = tuple(n[:nd])
    spacing = [float(_scalar(v, k)) for k in ("dx", "dy", "dz")[:nd]]
    kg = kWaveGrid(list(shape), spacing)
    kg.setTime(int(_scalar(v, "Nt")), float(_scalar(v, "dt")))
    absorbing = int(_scalar(v, "absorbing_flag", 0))
    med = kWaveMedium(sound_speed=_field(v, "c0", shape, nd), density=_field(v, "rho0", shape, nd))
    if absorbing:
        med.alpha_coeff = _field(v, "alpha_coeff", shape, nd)
        med.alpha_power = 2.0 if absorbing == 2 else float(_scalar(v, "alpha_power", 1.5))
    if int(_scalar(v, "nonlinear_flag", 0)):
        med.BonA = _field(v, "BonA", shape, nd)
    rho0_sg = [_field(v, f"rho0_sg{a}", shape, nd) for a in "xyz"[:nd]] if "rho0_sgx" in v else None
    from .cpp_simulation import SOURCE_MODE_NAME

    src = kSource()
    if int(_scalar(v, "p0_source_flag", 0)):
        src.p0 = _field(v, "p0_source_input", shape, nd)
    if int(_scalar(v, "p_source_flag", 0)):
        src.p_mask = _mask_from_index(v["p_source_index"], shape)
        src.p = _c_rows(_signals(v, "p_source_input", nd), v["p_source_index"], shape)
        src.p_mode = SOURCE_MODE_NAME[int(_scalar(v, "p_source_mode", 2))]
    if any(int(_scalar(v, f"{c}_source_flag", 0)) for c in ("ux", "uy", "uz")[:nd]):
        src.u_mask = _mask_from_index(v["u_source_index"], shape)
        src.u_mode = SOURCE_MODE_NAME[int(_scalar(v, "u_source_mode", 2))]
        for c in ("ux", "uy", "uz")[:nd]:
            if int(_scalar(v, f"{c}_source_flag", 0)):
                setattr(src, c, _c_rows(_signals(v, f"{c}_source_input", nd), v["u_source_index"], shape))
    if int(_scalar(v, "sensor_mask_type", 0)) != 0:
        raise ValueError("cuboid-corner sensor masks (sensor_mask_type 1) are not supported")
    sen = kSensor(_mask_from_index(v["sensor_mask_index"], shape))
    pml_size = tuple(int(_scalar(v, f"pml_{a}_size", 0)) for a in "xyz"[:nd])
    pml_alpha = tuple(float(_scalar(v, f"pml_{a}_alpha", 2.0)) for a in "xyz"[:nd])
    return kg, med, src, sen, pml_size, pml_alpha, rho0_sg


def run_input_file(input_file, output_file, *, device=0, record=("p",)):
    """Execute an input file on the GPU and write the output file; returns the output variables written."""
    from .. import _abi
    from .kspace_solver import _resolve_dtype, collect_results
    from .plan import DevicePlan
    from .setup import build_instance

    v = h5io.read_file(input_file)
    kg, med, src, sen, pml_size, pml_alpha, rho0_sg = problem_from_file(v)
    nd, shape = kg.dim, tuple(int(x) for x in kg.N)
    want = set(record)
    keys = set()
    if "p" in want:
        keys.add("p")
    keys.update(k for k in ("p_max", "p_min", "p_rms", "p_final") if k in want)
    if want & {"p_max_all", "p_min_all"}:
        keys.add("p")
    axes = "xyz"[:nd]
    if "u" in want:
        keys.update(f"u{a}_staggered" for a in axes)
    if "u_non_staggered" in want:
        keys.update(f"u{a}" for a in axes)
    if "u_final" in want:
        keys.update(f"u{a}_final" for a in axes)
    sen.record = sorted(keys)
    st = build_instance(kg, med, src, sen, pml_size=pml_size, pml_alpha=pml_alpha, smooth_p0=False,
                        source_scaled=True, rho0_sg=rho0_sg)
    plan = DevicePlan([st], device=device)
    try:
        plan.run(st.Nt - plan.step_index(0))
        res = collect_results(plan, 0, st, _resolve_dtype(np.float32))
        # the binaries write final fields over the whole grid, PML included (the caller crops them)
        if "p_final" in want:
            res["p_final"] = plan.field(0, _abi.F_P).copy()
        if "u_final" in want:
            for i, a in enumerate("xyz"[:nd]):
                res[f"u{a}_final"] = plan.field(0, _abi.F_UX + i).copy()
    finally:
        plan.close()
    # rows in Fortran order of the sensor points, as the binaries write them
    smask = np.asarray(sen.mask, dtype=bool)
    f = np.where(smask.ravel(order="F"))[0]
    ranks = np.argsort(np.argsort(np.ravel_multi_index(np.unravel_index(f, shape, order="F"), shape), kind="stable"))
    out = {}
    if "p" in want:
        out["p"] = res["p"][ranks]
    for k in ("p_max", "p_min", "p_rms"):
        if k in want:
            out[k] = res[k][ranks]
    if "p_final" in want:
        out["p_final"] = res["p_final"]
    if want & {"p_max_all", "p_min_all"}:
        raise ValueError("p_max_all / p_min_all need whole-grid reductions, which this runner does not record")
    for a in axes:
        if "u" in want:
            out[f"u{a}"] = res[f"u{a}_staggered"][ranks]
        if "u_non_staggered" in want:
            out[f"u{a}_non_staggered"] = res[f"u{a}"][ranks]
        if "u_final" in want:
            out[f"u{a}_final"] = res[f"u{a}_final"]
    for k, val in out.items():
        a = np.asarray(val, dtype=np.float32)
        # time series (points, steps) are stored so that reading and squeezing gives (points, steps)
        h5io.write_matrix(output_file, a.T if a.ndim == 2 and not k.endswith("_final") else a, k)
    h5io.write_attributes(output_file, file_type="output")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="Run a k-Wave C++-binary input file on the MI355X solver.")
    ap.add_argument("-i", dest="input", required=True)
    ap.add_argument("-o", dest="output", required=True)
    ap.add_argument("-g", dest="device", type=int, default=0)
    ap.add_argument("-t", dest="threads", type=int, default=None, help="accepted for compatibility; unused")
    ap.add_argument("--verbose", type=int, default=1)
    for f in RECORD_FLAGS:
        ap.add_argument(f"--{f}", action="store_true")
    args = ap.parse_args(argv)
    rec = {"p_raw": "p", "u_raw": "u", "u_non_staggered_raw": "u_non_staggered"}
    record = [rec.get(f, f) for f in RECORD_FLAGS if getattr(args, f)] or ["p"]
    out = run_input_file(args.input, args.output, device=args.device, record=record)
    if args.verbose:
        print(f"wrote {', '.join(sorted(out))} to {args.output}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
