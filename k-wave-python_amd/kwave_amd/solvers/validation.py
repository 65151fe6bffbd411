"""Pre-dispatch input checks (mirrors kwave/solvers/validation.py:8-108)."""
import warnings

import numpy as np


def validate_simulation(kgrid, medium, source, sensor, *, pml_size):
    validate_time_stepping(kgrid)
    validate_medium(medium, kgrid)
    validate_pml(pml_size, kgrid)
    validate_cfl(kgrid, medium)
    validate_source(source, kgrid)
    validate_sensor(sensor, kgrid)


def validate_time_stepping(kgrid):
    if kgrid.Nt == "auto" or kgrid.dt == "auto":
        raise ValueError("kgrid.Nt and kgrid.dt must be set before calling kspaceFirstOrder. "
                         "Call kgrid.makeTime(medium.sound_speed) or kgrid.setTime(Nt, dt) first.")
    if int(kgrid.Nt) < 1:
        raise ValueError(f"kgrid.Nt must be >= 1, got {kgrid.Nt}")
    if float(kgrid.dt) <= 0:
        raise ValueError(f"kgrid.dt must be > 0, got {kgrid.dt}")


def validate_medium(medium, kgrid):
    c = np.atleast_1d(np.asarray(medium.sound_speed))
    n = int(np.prod(kgrid.N))
    if c.size > 1 and c.size != n:
        raise ValueError(f"medium.sound_speed has {c.size} elements but grid has {n} points.")
    if np.any(c <= 0):
        raise ValueError("medium.sound_speed must be positive everywhere.")
    if medium.density is not None:
        rho = np.atleast_1d(np.asarray(medium.density))
        if rho.size > 1 and rho.size != n:
            raise ValueError(f"medium.density has {rho.size} elements but grid has {n} points.")
        if np.any(rho <= 0):
            raise ValueError("medium.density must be positive everywhere.")
    if medium.alpha_coeff is not None:
        if np.any(np.atleast_1d(np.asarray(medium.alpha_coeff)) < 0):
            raise ValueError("medium.alpha_coeff must be non-negative.")
        if medium.alpha_power is not None:
            y = float(np.asarray(medium.alpha_power).flat[0])
            if y < 0 or y > 3:
                warnings.warn(f"medium.alpha_power={y} is outside typical range [0, 3].", stacklevel=3)


def validate_pml(pml_size, kgrid):
    for i in range(kgrid.dim):
        N, pml = int(kgrid.N[i]), int(pml_size[i])
        if pml < 0:
            raise ValueError(f"pml_size[{i}] must be non-negative, got {pml}")
        if 2 * pml >= N:
            raise ValueError(f"pml_size[{i}]={pml} is too large for grid size N[{i}]={N} (2*pml must be < N).")


def validate_cfl(kgrid, medium):
    """Warn above the pseudospectral limit 2/(pi sqrt(D)) (Tabei et al., JASA 2002, eq. 11)."""
    cfl = float(np.max(np.asarray(medium.sound_speed))) * float(kgrid.dt) / float(np.min(kgrid.spacing))
    limit = 2 / (np.pi * np.sqrt(kgrid.dim))
    if cfl > limit:
        warnings.warn(f"CFL number = {cfl:.3f} exceeds pseudospectral stability limit {limit:.3f} for {kgrid.dim}D. "
                      "Reduce dt or increase grid spacing.", stacklevel=3)


def _is_cart(mask, dim):
    return mask.ndim == 2 and mask.shape[0] == dim


def validate_source(source, kgrid):
    n = int(np.prod(kgrid.N))
    if source.p0 is not None:
        p0 = np.asarray(source.p0)
        if p0.size != n and p0.size != 0:
            raise ValueError(f"source.p0 has {p0.size} elements but grid has {n} points.")
    if source.p is not None and source.p_mask is None:
        raise ValueError("source.p requires source.p_mask to be set.")
    for key in ("p_mask", "u_mask"):
        mask = getattr(source, key, None)
        if mask is not None:
            m = np.asarray(mask)
            if not _is_cart(m, kgrid.dim) and m.size != n:
                raise ValueError(f"source.{key} has {m.size} elements but grid has {n} points.")
    for vel in ("ux", "uy", "uz"):
        if getattr(source, vel, None) is not None and source.u_mask is None:
            raise ValueError(f"source.{vel} requires source.u_mask to be set.")
    dims = [np.squeeze(np.asarray(getattr(source, v))).ndim for v in ("ux", "uy", "uz") if getattr(source, v, None) is not None]
    if dims and len(set(d > 1 for d in dims)) > 1:
        raise ValueError("All velocity source components must either be single (1D) or many (2D) time series.")


def validate_sensor(sensor, kgrid):
    if sensor is None or getattr(sensor, "mask", None) is None:
        return
    m = np.asarray(sensor.mask)
    n = int(np.prod(kgrid.N))
    if not _is_cart(m, kgrid.dim) and m.size != 1 and m.size != n:
        raise ValueError(f"sensor.mask has {m.size} elements but grid has {n} points.")
