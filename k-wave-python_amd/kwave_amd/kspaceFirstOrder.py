"""Unified entry point (mirrors kwave/kspaceFirstOrder.py:16-350) on the GPU.

PML-outside expansion and stripping, dtype resolution, ``pml_size="auto"``,
validation and p0 smoothing after expansion, then the device ``Simulation``
(backend "python") or the C++ binaries' HDF5 file route run on the GPU
(backend "cpp", solvers/cpp_simulation.py).
"""
import copy
import warnings
from typing import Optional

import numpy as np

from .kgrid import kWaveGrid
from .utils.matrix import expand_matrix
from .utils.pml import get_optimal_pml_size


def _normalize_pml(val, ndim, name="pml_size"):
    """kspaceFirstOrder.py:16-25."""
    if isinstance(val, str):
        valid = "an integer, tuple of integers, or 'auto'" if name == "pml_size" else "a float or tuple of floats"
        raise ValueError(f"{name} must be {valid}, got {val!r}")
    if isinstance(val, (int, float, np.integer, np.floating)):
        return (val,) * ndim
    t = tuple(val)
    if len(t) == 0 or len(t) > ndim:
        raise ValueError(f"{name} must be a scalar or 1-to-{ndim}-element sequence, got {len(t)} elements")
    return t + (t[-1],) * (ndim - len(t))


def _is_cartesian_mask(mask, ndim):
    arr = np.asarray(mask)
    return arr.ndim == 2 and arr.shape[0] == ndim


def _expand_obj(obj, attrs, pml_size, edge_val=None):
    """Copy ``obj`` and pad its grid-sized attributes (kspaceFirstOrder.py:34-48)."""
    out = copy.copy(obj)
    for attr in attrs:
        val = getattr(out, attr, None)
        if val is not None:
            arr = np.asarray(val)
            if arr.size > 1:
                setattr(out, attr, expand_matrix(arr, list(pml_size), edge_val))
    return out


def _expand_for_pml_outside(kgrid, medium, source, sensor, pml_size):
    """Grow the grid by 2*pml per axis so the PML sits outside the user domain (kspaceFirstOrder.py:51-68)."""
    new_N = tuple(int(n) + 2 * int(p) for n, p in zip(kgrid.N, pml_size))
    kg = kWaveGrid(new_N, kgrid.spacing)
    kg.setTime(int(kgrid.Nt), kgrid.dt)
    med = _expand_obj(medium, ("sound_speed", "sound_speed_ref", "density", "alpha_coeff", "BonA"), pml_size)
    med = _expand_obj(med, ("alpha_filter",), pml_size, edge_val=0)
    src = _expand_obj(source, ("p0", "p_mask", "u_mask"), pml_size, edge_val=0)
    sen = sensor
    if sensor is not None and getattr(sensor, "mask", None) is not None and not _is_cartesian_mask(sensor.mask, kgrid.dim):
        sen = copy.copy(sensor)
        sen.mask = expand_matrix(np.asarray(sensor.mask), list(pml_size), 0)
    return kg, med, src, sen


_FULL_GRID_SUFFIXES = ("_max", "_min", "_rms", "_max_all", "_min_all", "_rms_all")


def _strip_pml(result, pml_size, ndim, suffixes=_FULL_GRID_SUFFIXES):
    """Crop full-grid outputs back to the user domain (kspaceFirstOrder.py:76-82)."""
    sl = tuple(slice(int(p), -int(p) if int(p) else None) for p in pml_size[:ndim])
    return {k: v[sl] if isinstance(v, np.ndarray) and v.ndim == ndim and any(k.endswith(s) for s in suffixes) else v
            for k, v in result.items()}


def _resolve_dtype(value):
    """kspaceFirstOrder.py:85-116."""
    if value is None or (isinstance(value, str) and value == "off"):
        return np.float64
    try:
        resolved = np.dtype(value).type
    except TypeError as e:
        framework = getattr(type(value), "__module__", "").split(".")[0]
        hint = ""
        if framework in ("torch", "jax", "jaxlib", "tensorflow"):
            hint = f" {framework}.dtype objects aren't supported; pass the equivalent numpy dtype (np.float32 / np.float64)."
        raise ValueError(f"dtype must be a numpy dtype, type, or string (e.g. 'float32', 'single'), got {value!r}.{hint}") from e
    if resolved is np.float32:
        return np.float32
    if resolved is np.float64:
        return np.float64
    raise ValueError(f"dtype must resolve to float32 or float64; got {resolved.__name__} from {value!r}")


def kspaceFirstOrder(kgrid, medium, source, sensor=None, *, pml_size=20, pml_alpha=2.0, pml_inside=False, use_sg=True,
                     use_kspace=True, smooth_p0=True, backend="python", device="gpu", dtype=None, save_only=False,
                     data_path: Optional[str] = None, quiet=False, debug=False, num_threads=None, device_num=None,
                     binary_path=None) -> dict:
    """Run a k-Wave simulation on the MI355X (kspaceFirstOrder.py:119-326).

    ``backend`` "python" is the native solver path; "cpp" writes the k-Wave
    C++ binaries' HDF5 input file and runs it on the MI355X (``save_only``
    writes the file only).  ``device`` "gpu" or "cpu" both run the HIP solver;
    ``device_num`` selects the HIP device.
    """
    if device not in ("cpu", "gpu"):
        raise ValueError(f"device must be 'cpu' or 'gpu', got {device!r}")
    if backend not in ("python", "cpp"):
        raise ValueError(f"Unknown backend: {backend!r}. Use 'python' or 'cpp'.")
    dtype = _resolve_dtype(dtype)
    if isinstance(pml_size, str) and pml_size.lower() == "auto":
        pml_size = tuple(int(x) for x in get_optimal_pml_size(kgrid))
    pml_size = _normalize_pml(pml_size, kgrid.dim)
    pml_alpha = _normalize_pml(pml_alpha, kgrid.dim, "pml_alpha")

    from .solvers.validation import validate_simulation

    validate_simulation(kgrid, medium, source, sensor, pml_size=pml_size)
    if not pml_inside:
        kgrid, medium, source, sensor = _expand_for_pml_outside(kgrid, medium, source, sensor, pml_size)
    if smooth_p0 and source.p0 is not None and kgrid.dim >= 2:
        from .utils.filters import smooth

        source = copy.copy(source)
        source.p0 = smooth(np.asarray(source.p0, dtype=float).reshape(tuple(int(n) for n in kgrid.N)), restore_max=True)
    from .solvers.kspace_solver import Simulation

    if backend == "cpp" and kgrid.dim == 1:
        raise ValueError("1-D grids run with backend='python' (the C++ binaries' files here are 2-D and 3-D)")
    if backend == "cpp":
        result = _run_cpp_route(kgrid, medium, source, sensor, pml_size, pml_alpha, use_sg, use_kspace, dtype, save_only,
                                data_path, device_num)
        if save_only:
            if not pml_inside:
                result["pml_size"] = pml_size
            return result
        if not pml_inside:
            result = _strip_pml(result, pml_size, kgrid.dim, suffixes=("_final",) + _FULL_GRID_SUFFIXES)
        return result
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", UserWarning)
        result = Simulation(kgrid, medium, source, sensor, device=device_num if device_num is not None else "gpu",
                            use_sg=use_sg, use_kspace=use_kspace, smooth_p0=False, pml_size=pml_size,
                            pml_alpha=pml_alpha, quiet=quiet, dtype=dtype).run()
    if not pml_inside:
        result = _strip_pml(result, pml_size, kgrid.dim)
    return result


_CPP_UNSUPPORTED_ALPHA_MODES = ("no_absorption", "no_dispersion", "stokes")
_DISPERSION_SINGULAR = (0.95, 1.05)


def _run_cpp_route(kgrid, medium, source, sensor, pml_size, pml_alpha, use_sg, use_kspace, dtype, save_only, data_path,
                   device_num):
    """backend="cpp": the C++ binaries' file route (kspaceFirstOrder.py:271-319), run on the MI355X
    (solvers/cpp_simulation.py).  Same checks and warnings as the reference's route."""
    from .solvers.cpp_simulation import CppSimulation
    from .utils.conversion import cart2grid

    mode = getattr(medium, "alpha_mode", None)
    mode = getattr(mode, "value", mode)
    if mode in _CPP_UNSUPPORTED_ALPHA_MODES:
        raise ValueError(f"medium.alpha_mode={mode!r} is not supported by the C++ backend: its HDF5 input format does not "
                         "carry alpha_mode, so full power-law absorption and dispersion would apply. Use backend='python'.")
    a, y = getattr(medium, "alpha_coeff", None), getattr(medium, "alpha_power", None)
    if a is not None and y is not None and np.any(np.asarray(a) > 0):
        yy = np.asarray(y, dtype=float)
        if np.any((yy >= _DISPERSION_SINGULAR[0]) & (yy <= _DISPERSION_SINGULAR[1])):
            warnings.warn(f"medium.alpha_power in the dispersion-singular range {list(_DISPERSION_SINGULAR)} on the C++ backend: "
                          "tan(pi y / 2) overflows in float32 (issue 664); use backend='python' with "
                          "alpha_mode='no_dispersion'.", stacklevel=3)
    if dtype is not np.float64:
        warnings.warn(f"dtype={np.dtype(dtype).name!r} has no effect with backend='cpp'.", stacklevel=3)
    if not use_kspace:
        warnings.warn("use_kspace=False has no effect with backend='cpp'; the k-space correction always applies.", stacklevel=3)
    if sensor is not None and getattr(sensor, "record", None) is not None:
        warnings.warn("sensor.record is not yet supported for backend='cpp'; the pressure time series is recorded.",
                      stacklevel=3)
    if sensor is not None and getattr(sensor, "record_start_index", 1) != 1:
        warnings.warn("sensor.record_start_index is not yet supported for backend='cpp'; recording starts at the first step.",
                      stacklevel=3)
    if sensor is not None and getattr(sensor, "mask", None) is not None and _is_cartesian_mask(sensor.mask, kgrid.dim):
        sensor = copy.copy(sensor)
        sensor.mask = cart2grid(kgrid, np.asarray(sensor.mask), order="C")[0] != 0
    sim = CppSimulation(kgrid, medium, source, sensor, pml_size=pml_size, pml_alpha=pml_alpha, use_sg=use_sg)
    if save_only:
        if data_path is None:
            raise ValueError("data_path must be provided when save_only=True (the HDF5 files must persist for cluster submission).")
        inp, outp = sim.prepare(data_path=data_path)
        return {"input_file": inp, "output_file": outp}
    return sim.run(device_num=device_num, data_path=data_path)


def reshape_to_grid(data, grid_shape):
    """(n_sensor, Nt) -> (*grid, Nt) for full-grid masks (kspaceFirstOrder.py:329-350)."""
    data = np.asarray(data)
    if data.ndim == 2:
        return data.reshape(*grid_shape, data.shape[1])
    if data.ndim == 1:
        return data.reshape(grid_shape)
    return data
