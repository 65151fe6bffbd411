"""1-D simulations on the 2-D kernels.

The reference's Simulation runs 1-D grids (kspace_solver.py:343-350 for Cartesian
sensors, :682 for the p0 smoothing it skips in 1-D).  Here a 1-D problem of N
points runs as a 2 x N grid that is uniform along axis 0, the 1-D axis being the
contiguous axis 1: every field is the same in both rows, so the axis-0 derivative
is zero and the sum of the split densities obeys the 1-D equations.  Two choices
keep that sum identical to the 1-D density everywhere, PML included: axis 0 has no
PML, and everything the 1-D solver puts into its single density - the initial
density p0 / c0^2 and the pressure-source terms - goes into the axis-1 split
density (kw_plan_set_density_split shares {0, 1}; pressure-source factors
{0, 2 / dx} and {0, 2} against the 2-D scale that divides by D = 2).  Results are
read back from row 0, with the axis-1 velocity and intensity reported under the
1-D names (ux, Ix).
"""
from __future__ import annotations

import copy

import numpy as np

ROW = 0  # the row whose values are returned


def _ndim(kgrid):
    if hasattr(kgrid, "N"):
        return int(np.atleast_1d(kgrid.N).size)
    return int(getattr(kgrid, "dim", 0))


def is_1d(kgrid):
    return _ndim(kgrid) == 1


def _rows(a, n):
    """A per-point 1-D array as the 2 x n grid (both rows); scalars stay scalars."""
    if a is None:
        return None
    arr = np.asarray(a)
    if arr.size <= 1:
        return a
    if arr.size != n:
        raise ValueError(f"1-D field has {arr.size} values for a grid of {n} points")
    return np.repeat(arr.reshape(1, n), 2, axis=0)


def _signal_rows(sig, npts):
    """Signals of a mask that is duplicated into both rows: one row stays one row, one row per
    point becomes the rows of row 0's points followed by the same rows for row 1's points."""
    if sig is None:
        return None
    s = np.asarray(sig, dtype=np.float64)
    if s.ndim == 1 or s.shape[0] == 1 or npts <= 1:
        return s
    return np.concatenate([s, s], axis=0)


_MAP_RECORD = {"u": "uy", "u_staggered": "uy_staggered", "u_non_staggered": "uy", "I": "Iy", "I_avg": "Iy_avg"}


def record_2d(record):
    """1-D record names -> the 2-D names of the propagation axis (axis 1)."""
    if record is None:
        return None
    rec = (record,) if isinstance(record, str) else tuple(record)
    out = []
    for k in rec:
        if k in _MAP_RECORD:
            out.append(_MAP_RECORD[k])
        elif k.startswith("ux"):
            out.append("uy" + k[2:])
        elif k.startswith("u_"):
            out.append("uy" + k[1:])
        elif k.startswith("Ix"):
            out.append("Iy" + k[2:])
        else:
            out.append(k)
    return out


def embed(kgrid, medium, source, sensor, pml_size, pml_alpha):
    """The 2 x N problem of a 1-D problem: (kgrid, medium, source, sensor, pml_size, pml_alpha)."""
    from ..kgrid import kWaveGrid
    from ..ksensor import kSensor

    n = int(np.atleast_1d(kgrid.N)[0])
    dx = float(np.atleast_1d(kgrid.spacing)[0])
    kg = kWaveGrid([2, n], [dx, dx])
    kg.setTime(int(kgrid.Nt), float(kgrid.dt))
    med = copy.copy(medium)
    for attr in ("sound_speed", "density", "alpha_coeff", "BonA", "alpha_filter"):
        val = getattr(med, attr, None)
        if val is not None and np.asarray(val).size > 1:
            setattr(med, attr, _rows(val, n))
    src = copy.copy(source)
    if getattr(source, "p0", None) is not None:
        src.p0 = _rows(source.p0, n)
    for mname, signames in (("p_mask", ("p",)), ("u_mask", ("ux",))):
        mask = getattr(source, mname, None)
        if mask is None or np.asarray(mask).size <= 1:
            continue
        m1 = np.asarray(mask).astype(bool).ravel()
        setattr(src, mname, _rows(m1, n))
        for sname in signames:
            sig = getattr(source, sname, None)
            if sig is not None:
                setattr(src, sname, _signal_rows(sig, int(m1.sum())))
    if getattr(source, "u_mask", None) is not None:
        src.uy, src.ux = src.ux, None
    sen = sensor
    if sensor is not None:
        sen = copy.copy(sensor)
        # binary masks go in flattened: a 2 x n array would read as n Cartesian points in 2-D
        mask = getattr(sensor, "mask", None)
        if mask is None:
            m2 = np.zeros((2, n), dtype=bool)
            m2[ROW] = True
            sen.mask = m2.ravel()
        else:
            m = np.asarray(mask, dtype=float)
            if m.ndim == 2 and m.shape[0] == 1:  # Cartesian positions along the axis
                x0 = (ROW - 1) * dx                # axis-0 coordinate of row 0: (i - N0 // 2) d
                sen.mask = np.vstack([np.full((1, m.shape[1]), x0), m])
            else:
                m1 = np.asarray(mask).astype(bool).ravel()
                if m1.size == 1:
                    m1 = np.full(n, bool(m1[0]))
                m2 = np.zeros((2, n), dtype=bool)
                m2[ROW] = m1
                sen.mask = m2.ravel()
        sen.record = record_2d(getattr(sensor, "record", None))
    else:
        m2 = np.zeros((2, n), dtype=bool)
        m2[ROW] = True
        sen = kSensor(m2.ravel())
    pml = (0, int(np.atleast_1d(pml_size)[0]) if pml_size is not None else 0)
    alpha = (0.0, float(np.atleast_1d(pml_alpha)[0]) if pml_alpha is not None else 0.0)
    return kg, med, src, sen, pml, alpha


def adjust_setup(st):
    """Put the 1-D solver's whole density (p0 and pressure sources) into the axis-1 split density."""
    st.p0_split = (0.0, 1.0, 0.0)
    if st.src_p is not None:
        f = tuple(st.src_p.fac)
        st.src_p.fac = (0.0, 2.0 * f[1]) + tuple(f[2:])
    return st


def unembed(result):
    """Result dict of the 2 x N run -> the 1-D result dict (row 0, 1-D names)."""
    out = {}
    for k, v in result.items():
        if k.startswith("ux") or k.startswith("Ix"):
            continue  # axis-0 components of the embedding (zero)
        name = "ux" + k[2:] if k.startswith("uy") else "Ix" + k[2:] if k.startswith("Iy") else k
        arr = v
        if isinstance(v, np.ndarray) and v.ndim == 2 and v.shape[0] == 2 and (k.endswith("_final") or k.endswith("_all")):
            arr = v[ROW].copy()
        out[name] = arr
    return out
