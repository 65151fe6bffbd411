"""Many independent simulations of uneven grid sizes and step counts in one go.

Instances are grouped by grid shape (and physics switches) into device plans;
each plan holds its instances as a batch (one kernel launch per step covers
all of them, instances that reached their own Nt drop out).  All plans then run
as one kw_batch: every block of steps is a single captured hipGraph in which
each size group's step sequence is a parallel branch, so the groups execute
concurrently from one graph launch.

This is the batched counterpart of calling the native solver once per
instance (kwave/solvers/kspace_solver.py:749-778 in a Python loop).
"""
from __future__ import annotations

import ctypes as C
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

    def run(self, graphs="batch"):
        """Advance every instance to its own Nt; returns the per-instance result dicts in input order.

        ``graphs="batch"`` (default): kw_batch_run, one captured graph per block of steps with the
        size groups as parallel branches.  ``graphs="plans"``: every plan replays its own step
        graphs on its own stream (kw_plan_launch), all plans enqueued before any is waited for.
        Both give the same bits; the second keeps each graph to one stream (for tools that
        trace graph launches stream by stream).
        """
        if graphs not in ("batch", "plans"):
            raise ValueError(f"graphs must be 'batch' or 'plans', got {graphs!r}")
        if graphs == "plans":
            for plan, idx in zip(self.plans, self.groups):
                plan.launch(max(self.setups[i].Nt for i in idx) - min(plan.step_index(b) for b in range(len(idx))))
            for plan in self.plans:
                plan.sync()
        else:
            L = _abi.lib()
            handles = (C.c_void_p * len(self.plans))(*[p._h.value for p in self.plans])
            h = C.c_void_p()
            _abi.check(L.kw_batch_create(handles, len(self.plans), C.byref(h)))
            try:
                _abi.check(L.kw_batch_run(h))
            finally:
                L.kw_batch_destroy(h)
        out = [None] * len(self.setups)
        for plan, idx in zip(self.plans, self.groups):
            for b, i in enumerate(idx):
                out[i] = collect_results(plan, b, self.setups[i], self._dtype)
        return out

    def close(self):
        for p in self.plans:
            p.close()
        self.plans = []


def _main():
    """`python -m kwave_amd.batch [--count N] [--seed S] [--repeat R] [--graphs batch|plans]`:
    run the uneven 2-D scan-line sweep (workloads.scan_sweep_2d) batched and print the time
    of each run (the first includes graph capture)."""
    import argparse
    import time

    import numpy as np

    from . import workloads

    ap = argparse.ArgumentParser(description=_main.__doc__)
    ap.add_argument("--count", type=int, default=96)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--repeat", type=int, default=2)
    ap.add_argument("--graphs", choices=("batch", "plans"), default="batch")
    args = ap.parse_args()
    cases = workloads.scan_sweep_2d(seed=args.seed, count=args.count)
    for r in range(args.repeat):
        bs = BatchSimulation(cases)
        vs = sum(int(np.prod(s.shape)) * s.Nt for s in bs.setups)
        t0 = time.perf_counter()
        bs.run(graphs=args.graphs)
        dt = time.perf_counter() - t0
        bs.close()
        print(f"sweep run {r + 1}: {args.count} instances in {len(bs.groups)} size groups, {dt:.3f} s, "
              f"{vs / dt:.3e} voxel-steps/s", flush=True)


if __name__ == "__main__":
    _main()
