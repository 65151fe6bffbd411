"""Many independent simulations of uneven grid sizes and step counts in one go.

Instances are grouped by grid shape (and physics switches) into device plans;
each plan holds its instances as a batch (one kernel launch per step covers
all of them, instances that reached their own Nt drop out), and every plan
runs on its own stream, all enqueued before any synchronisation, so the size
groups execute concurrently.  Each plan replays a captured hipGraph per step.

This is the batched counterpart of calling the native solver once per
instance (kwave/solvers/kspace_solver.py:749-778 in a Python loop).
"""
from __future__ import annotations

from collections import OrderedDict

from . import _abi
from .solvers.kspace_solver import _resolve_dtype, collect_results
from .solvers.plan import DevicePlan
from .solvers.setup import build_instance


class BatchSimulation:
    """Run ``cases`` = [(kgrid, medium, source, sensor, pml_size, pml_alpha), ...] together."""

    def __init__(self, cases, *, device=0, use_sg=True, use_kspace=True, smooth_p0=True, dtype=None):
        self._dtype = _resolve_dtype(dtype)
        self.setups = [build_instance(kg, med, src, sen, use_sg=use_sg, use_kspace=use_kspace, smooth_p0=smooth_p0,
                                      pml_size=pml, pml_alpha=alpha) for kg, med, src, sen, pml, alpha in cases]
        groups = OrderedDict()
        for i, st in enumerate(self.setups):
            # the power-law exponent is a per-plan constant of the fractional Laplacians
            y = st.alpha_power if st.flags & (_abi.KW_STOKES | _abi.KW_POWERLAW_ABS | _abi.KW_POWERLAW_DISP) else None
            key = (st.shape, tuple(st.spacing), st.flags, st.use_sg, st.p0 is not None, y)
            groups.setdefault(key, []).append(i)
        self.groups = list(groups.values())
        self.plans = [DevicePlan([self.setups[i] for i in idx], device=device) for idx in self.groups]

    def run(self):
        """Advance every instance to its own Nt; returns the per-instance result dicts in input order."""
        for plan, idx in zip(self.plans, self.groups):
            plan.launch(max(self.setups[i].Nt for i in idx) - min(plan.step_index(b) for b in range(len(idx))))
        for plan in self.plans:
            plan.sync()
        out = [None] * len(self.setups)
        for plan, idx in zip(self.plans, self.groups):
            for b, i in enumerate(idx):
                out[i] = collect_results(plan, b, self.setups[i], self._dtype)
        return out

    def close(self):
        for p in self.plans:
            p.close()
        self.plans = []
