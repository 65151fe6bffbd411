"""Input checks run before a simulation is dispatched to the device.

Behavioural counterpart of kwave/solvers/validation.py:8-106: the same public
entry points (``validate_simulation`` and one function per input group), the
same exception types (ValueError for inputs the solver cannot run, a
UserWarning for a time step above the pseudospectral stability limit) and the
keywords the reference's tests match (tests/test_validation.py:31-104).

The checks are data: each group is a tuple of ``Rule``s, a rule is a named
predicate over a small context dict plus the message used when it fails.  A
group function builds the context and evaluates its rules in order, so the
first violated rule reports.
"""
from __future__ import annotations

import warnings
from typing import Callable, NamedTuple

import numpy as np


class Rule(NamedTuple):
    ok: Callable[[dict], bool]        # True when the input passes
    message: Callable[[dict], str]    # text of the failure
    warn: bool = False                # report with a warning instead of raising


def _check(rules, ctx):
    for rule in rules:
        if rule.ok(ctx):
            continue
        if rule.warn:
            warnings.warn(rule.message(ctx), UserWarning, stacklevel=4)
        else:
            raise ValueError(rule.message(ctx))


def _numel(kgrid):
    return int(np.prod(np.atleast_1d(kgrid.N)))


def _arr(x):
    return np.atleast_1d(np.asarray(x))


def _grid_sized_or_scalar(a, n):
    return a.size in (1, n)


def _cartesian(mask, dim):
    """A (dim, n_points) float array of coordinates rather than a binary grid mask."""
    return mask.ndim == 2 and mask.shape[0] == dim


# ---------------------------------------------------------------------------
# time stepping (kwave/solvers/validation.py time checks)
# ---------------------------------------------------------------------------
_TIME = (
    Rule(lambda c: not (isinstance(c["Nt"], str) or isinstance(c["dt"], str)),
         lambda c: "kgrid.Nt and kgrid.dt must be set before the simulation runs: call "
                   "kgrid.makeTime(medium.sound_speed) or kgrid.setTime(Nt, dt) first."),
    Rule(lambda c: int(c["Nt"]) >= 1, lambda c: f"kgrid.Nt must be at least 1 (got {c['Nt']})"),
    Rule(lambda c: float(c["dt"]) > 0, lambda c: f"kgrid.dt must be greater than zero (got {c['dt']})"),
)


def validate_time_stepping(kgrid):
    _check(_TIME, {"Nt": kgrid.Nt, "dt": kgrid.dt})


# ---------------------------------------------------------------------------
# medium
# ---------------------------------------------------------------------------
def _medium_rules(name):
    """Size and positivity of one medium property (sound_speed or density)."""
    return (
        Rule(lambda c: _grid_sized_or_scalar(c[name], c["n"]),
             lambda c: f"medium.{name}: {c[name].size} values given for a grid of {c['n']} points "
                       "(expected a scalar or one value per grid point)"),
        Rule(lambda c: bool(np.all(c[name] > 0)), lambda c: f"medium.{name} must be positive at every grid point"),
    )


_ALPHA = (
    Rule(lambda c: c["alpha"] is None or bool(np.all(c["alpha"] >= 0)),
         lambda c: "medium.alpha_coeff must be non-negative"),
    Rule(lambda c: c["y"] is None or 0 <= c["y"] <= 3,
         lambda c: f"medium.alpha_power = {c['y']} lies outside the usual power-law range [0, 3]", warn=True),
)


def validate_medium(medium, kgrid):
    n = _numel(kgrid)
    ctx = {"n": n, "sound_speed": _arr(medium.sound_speed).astype(float)}
    _check(_medium_rules("sound_speed"), ctx)
    if getattr(medium, "density", None) is not None:
        ctx["density"] = _arr(medium.density).astype(float)
        _check(_medium_rules("density"), ctx)
    alpha = getattr(medium, "alpha_coeff", None)
    power = getattr(medium, "alpha_power", None)
    ctx["alpha"] = None if alpha is None else _arr(alpha).astype(float)
    ctx["y"] = None if (alpha is None or power is None) else float(_arr(power).flat[0])
    _check(_ALPHA, ctx)


# ---------------------------------------------------------------------------
# PML
# ---------------------------------------------------------------------------
_PML = (
    Rule(lambda c: c["size"] >= 0, lambda c: f"pml_size[{c['axis']}] must be non-negative (got {c['size']})"),
    Rule(lambda c: 2 * c["size"] < c["N"],
         lambda c: f"pml_size[{c['axis']}] = {c['size']} is too large for an axis of {c['N']} points "
                   "(the two layers must leave an interior: 2 * pml_size < N)"),
)


def validate_pml(pml_size, kgrid):
    N = np.atleast_1d(kgrid.N)
    for axis in range(int(kgrid.dim)):
        _check(_PML, {"axis": axis, "size": int(pml_size[axis]), "N": int(N[axis])})


# ---------------------------------------------------------------------------
# CFL
# ---------------------------------------------------------------------------
# Stability of the leapfrog pseudospectral scheme on a heterogeneous grid:
# c_max dt / dx_min <= 2 / (pi sqrt(D)) (Tabei, Mast and Waag, JASA 111, 2002).
_CFL = (
    Rule(lambda c: c["cfl"] <= c["limit"],
         lambda c: f"CFL number {c['cfl']:.3f} is above the pseudospectral stability limit {c['limit']:.3f} "
                   f"of a {c['dim']}-D grid; use a smaller dt or a coarser grid", warn=True),
)


def validate_cfl(kgrid, medium):
    dim = int(kgrid.dim)
    cfl = float(np.max(_arr(medium.sound_speed))) * float(kgrid.dt) / float(np.min(_arr(kgrid.spacing)))
    _check(_CFL, {"cfl": cfl, "limit": 2.0 / (np.pi * np.sqrt(dim)), "dim": dim})


# ---------------------------------------------------------------------------
# sources
# ---------------------------------------------------------------------------
def _mask_fits(c, key):
    m = c[key]
    return m is None or _cartesian(m, c["dim"]) or m.size == c["n"]


def _series_kinds(c):
    """1 per velocity component given as one time series, 2 per component given per source point."""
    return {int(np.squeeze(v).ndim > 1) for v in c["u_series"]}


_SOURCE = (
    Rule(lambda c: c["p0"] is None or c["p0"].size in (0, c["n"]),
         lambda c: f"source.p0: {c['p0'].size} values given for a grid of {c['n']} points"),
    Rule(lambda c: not (c["has_p"] and c["p_mask"] is None),
         lambda c: "a pressure source (source.p) needs source.p_mask"),
    Rule(lambda c: _mask_fits(c, "p_mask"),
         lambda c: f"source.p_mask: {c['p_mask'].size} values given for a grid of {c['n']} points"),
    Rule(lambda c: _mask_fits(c, "u_mask"),
         lambda c: f"source.u_mask: {c['u_mask'].size} values given for a grid of {c['n']} points"),
    Rule(lambda c: not (c["u_series"] and c["u_mask"] is None),
         lambda c: f"velocity sources ({', '.join('source.' + k for k in c['u_names'])}) need source.u_mask"),
    Rule(lambda c: len(_series_kinds(c)) <= 1,
         lambda c: "velocity source components must all be single time series or all be one series per "
                   "source point (mixed 1-D and 2-D inputs)"),
)


def validate_source(source, kgrid):
    def opt(name):
        v = getattr(source, name, None)
        return None if v is None else np.asarray(v)

    names = [k for k in ("ux", "uy", "uz") if getattr(source, k, None) is not None]
    ctx = {"n": _numel(kgrid), "dim": int(kgrid.dim), "p0": opt("p0"), "has_p": getattr(source, "p", None) is not None,
           "p_mask": opt("p_mask"), "u_mask": opt("u_mask"), "u_names": names,
           "u_series": [np.asarray(getattr(source, k)) for k in names]}
    _check(_SOURCE, ctx)


# ---------------------------------------------------------------------------
# sensor
# ---------------------------------------------------------------------------
_SENSOR = (
    Rule(lambda c: c["mask"] is None or _cartesian(c["mask"], c["dim"]) or c["mask"].size in (1, c["n"]),
         lambda c: f"sensor.mask: {c['mask'].size} values given for a grid of {c['n']} points "
                   "(expected a binary grid mask or Cartesian points of shape (dim, n))"),
)


def validate_sensor(sensor, kgrid):
    mask = None if sensor is None else getattr(sensor, "mask", None)
    _check(_SENSOR, {"mask": None if mask is None else np.asarray(mask), "n": _numel(kgrid), "dim": int(kgrid.dim)})


def validate_simulation(kgrid, medium, source, sensor, *, pml_size):
    """Every group, time stepping first (the CFL check needs dt)."""
    validate_time_stepping(kgrid)
    validate_medium(medium, kgrid)
    validate_pml(pml_size, kgrid)
    validate_cfl(kgrid, medium)
    validate_source(source, kgrid)
    validate_sensor(sensor, kgrid)
