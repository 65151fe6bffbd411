"""Gradient of recorded pressure traces with respect to sound speed (reverse-time adjoint on the GPU).

The reference solver (kwave/solvers/kspace_solver.py) is forward-only.  This
module differentiates its step (``Simulation.step``, :691-747) for lossless
media, media with Stokes or power-law absorption and dispersion (:471-516; tau
and eta depend on c0) and B/A nonlinearity (:542-550), with or without absorption,
with additive or dirichlet pressure and velocity sources and/or an initial pressure p0 (whose t = 0 state,
:729-733, depends on c0 through the initial densities p0 / (D c0^2)) and binary or
Cartesian sensors, in 1-D, 2-D and 3-D: ``sound_speed_gradient`` returns dJ/dc0 for a given dJ/d(traces), and
``sensor_traces`` wraps forward + adjoint as a ``torch.autograd.Function`` so a
misfit written in torch can be minimised over the sound-speed map.

The forward state is checkpointed every ``checkpoint`` steps and replayed one
segment at a time (memory: (2D+1) grids per checkpoint plus ``checkpoint``
pressure grids; absorbing media keep one or two more grids per replayed step, B/A
one more grid per checkpoint and D+1 more per replayed step).  c_ref = max(c0) (kspace_solver.py:231), dt, the PML and kappa
are held fixed: the gradient is exact for perturbations that keep max(c0).
"""
from __future__ import annotations

import copy

import numpy as np

from . import _abi
from .solvers import embed1d
from .solvers.plan import DevicePlan
from .solvers.setup import build_instance


def _plan(kgrid, medium, source, sensor, c, pml_size, pml_alpha, **kw):
    """(device plan, set-up, 1-D?).  A 1-D problem runs as the 2 x N grid of solvers/embed1d.py,
    c0 replicated into both rows."""
    if kw.get("source_scaled"):  # pre-scaled signals carry no c0 dependence the adjoint could use
        raise ValueError("adjoint: source_scaled signals are not differentiated")
    med = copy.copy(medium)
    if c is not None:
        med.sound_speed = np.asarray(c, dtype=np.float64)
    one_d = embed1d.is_1d(kgrid)
    if one_d:
        kgrid, med, source, sensor, pml_size, pml_alpha = embed1d.embed(kgrid, med, source, sensor, pml_size, pml_alpha)
        kw = dict(kw, smooth_p0=False)  # the reference does not smooth p0 in 1-D (kspace_solver.py:682)
    st = build_instance(kgrid, med, copy.deepcopy(source), sensor, pml_size=pml_size, pml_alpha=pml_alpha, **kw)
    if one_d:
        st = embed1d.adjust_setup(st)
    if "p" not in st.record:
        raise ValueError("the sensor must record 'p'")
    return DevicePlan([st]), st, one_d


def forward_traces(kgrid, medium, source, sensor, *, c=None, pml_size=None, pml_alpha=None, **kw):
    """Recorded p of one simulation, (n_sensor, n_recorded_steps) float32."""
    plan, st, _ = _plan(kgrid, medium, source, sensor, c, pml_size, pml_alpha, **kw)
    try:
        plan.run(st.Nt - plan.step_index(0))
        return plan.recorded(0, _abi.F_REC_P).copy()
    finally:
        plan.close()


def sound_speed_gradient(kgrid, medium, source, sensor, adj_src, *, c=None, pml_size=None, pml_alpha=None,
                         checkpoint=50, **kw):
    """(dJ/dc0 as a grid, recorded p) for dJ/d(recorded p) = ``adj_src`` (n_sensor, n_rec)."""
    plan, st, one_d = _plan(kgrid, medium, source, sensor, c, pml_size, pml_alpha, **kw)
    try:
        grad = plan.adjoint_c([adj_src], checkpoint=checkpoint)[0]
        if one_d:  # c0 of a 1-D point sets both rows of the 2 x N grid
            grad = grad.sum(axis=0)
        return grad, plan.recorded(0, _abi.F_REC_P).copy()
    finally:
        plan.close()


def _autograd():
    import torch

    if torch.cuda.is_available():
        torch.cuda.init()  # PyTorch's HIP context before the solver library's (see _abi._torch_context_first)

    class SensorTraces(torch.autograd.Function):
        """c0 map -> recorded p traces; backward runs the reverse-time adjoint."""

        @staticmethod
        def forward(ctx, c, problem):
            c_np = c.detach().to("cpu", torch.float64).numpy()
            kg, med, src, sen, kw = problem
            kw = {k: v for k, v in kw.items() if k != "checkpoint"}
            traces = forward_traces(kg, med, src, sen, c=c_np, **kw)
            ctx.problem, ctx.c = problem, c_np
            ctx.dtype, ctx.device = c.dtype, c.device
            return torch.from_numpy(traces).to(device=c.device, dtype=torch.float32)

        @staticmethod
        def backward(ctx, g):
            kg, med, src, sen, kw = ctx.problem
            adj = g.detach().to("cpu", torch.float32).numpy()
            grad, _ = sound_speed_gradient(kg, med, src, sen, adj, c=ctx.c, **kw)
            return torch.from_numpy(grad).to(device=ctx.device, dtype=ctx.dtype), None

    return SensorTraces


_SensorTraces = None


def sensor_traces(c, kgrid, medium, source, sensor, **kw):
    """Differentiable recorded p (torch tensor, (n_sensor, n_rec)) as a function of the c0 tensor ``c``.

    Keywords go to the set-up (``pml_size``, ``pml_alpha``, ``checkpoint`` ...).
    """
    global _SensorTraces
    if _SensorTraces is None:
        _SensorTraces = _autograd()
    return _SensorTraces.apply(c, (kgrid, medium, source, sensor, kw))
