#!/usr/bin/env python3
"""Benchmark: the 512^3 heterogeneous nonlinear absorbing k-space PSTD time loop on one MI355X.

Workload (frozen in kwave_amd/workloads.py::nonlinear_3d): 512^3 grid with a
16-cell PML, Gaussian-random c in [1450, 1650] m/s and rho in [950, 1100]
kg/m^3, B/A = 6, power-law absorption 0.5 dB/(MHz^y cm) with y = 1.1, a 16x16
planar additive source (0.5 MHz, 4 cycles, 2 MPa) and a 256x256 planar sensor
recording p and p_max, 1500 steps, float32 arithmetic.

A "step" is one time step of that simulation.  W warm-up steps run first, the
state is re-initialised, then exactly K steps are timed between two device
synchronisations, with HIP events bracketing every kernel launch on the plan
stream (per-kernel device time, used for the roofline of the dominant kernel).

Further legs (skippable): the two 2-D workloads (128^2 IVP, 512^2 absorbing),
the 96-instance uneven 2-D sweep batched vs looped, the learned surrogate
(trained on 4096 solver samples of the 128^2 sound-speed -> traces problem,
held-out error and speed against the batched solver) and the reverse-time
adjoint (gradient time against forward time), and the FP64 CPU oracle as the
CPU baseline.

Prints ONE JSON line on stdout.  Everything else goes to stderr.
"""
import argparse
import json
import math
import os
import platform
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "k-wave-python_amd"))

HBM_PEAK_GBS = 8000.0
# fused-minimum algorithmic bytes per voxel-step of this workload: 412 B
# (3-D transforms + fused pointwise, power-law, nonlinear) plus 60 B on the
# source's ~440 active steps of 1500
STEP_MODEL_BYTES = 412.0 + 60.0 * 440.0 / 1500.0


PROFILE = ROOT / "profiles" / "rocprof_3d_nonlinear_512.md"
# kernel kind (per-kernel HIP-event timing) -> compiled kernel name in the rocprof summary
ROCPROF_NAME = {"col_y": "kw::k_col<{n}, 16, false, 1, 0>", "z_mass": "kw::k_mass<{n}, 3, false, false>",
                "z_momentum": "kw::k_momentum<{n}>", "z_eos": "kw::k_eos<{n}>"}


def measured_traffic(kind, n):
    """Calibrated HBM bytes per dispatch from the committed FETCH_SIZE/WRITE_SIZE summary, or None."""
    name = ROCPROF_NAME.get(kind, "").format(n=n)
    if not name or not PROFILE.exists():
        return None
    for line in PROFILE.read_text().splitlines():
        if line.startswith("traffic ") and line[len("traffic "):].rsplit(" ", 1)[0] == name:
            return float(line.rsplit(" ", 1)[1])
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(n=256, n_target=512):
    """FP64 oracle on the same workload at n^3: one warm-up and one timed step, single-threaded
    numpy.fft (the reference's path), then the same with scipy.fft on every host core (at most 16).
    The step time is scaled to n_target^3 by N log N (N voxels; the step is FFT-dominated), so the
    rate scales by (N_target / N) / scale = log N / log N_target."""
    sys.path.insert(0, str(ROOT))
    from kwave_amd import workloads
    from oracle.kspace_ref import OracleSimulation

    kg, med, src, sen, pml, alpha = workloads.nonlinear_3d(seed=3, n=n, nt=4, pml=max(2, 16 * n // 512))
    scale = (n_target ** 3 * math.log(n_target ** 3)) / (n ** 3 * math.log(n ** 3))
    out = {}
    ncores = min(os.cpu_count() or 1, 16)
    for workers in (1, ncores):
        sim = OracleSimulation(kg, med, src, sen, pml_size=pml, pml_alpha=alpha, fft_workers=workers)
        sim.step()  # warm-up
        t0 = time.perf_counter()
        sim.step()
        dt = time.perf_counter() - t0
        out[workers] = (n ** 3 / dt, dt)
        del sim
        log(f"cpu baseline: {n}^3 step {dt:.2f} s with {workers} FFT thread(s)")
    rate1, dt1 = out[1]
    ratem, dtm = out[ncores]
    return {
        "value": rate1 * (n_target ** 3 / n ** 3) / scale,
        "unit": "voxel-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"FP64 numpy oracle (numpy.fft, as the reference), same 512^3 workload built at {n}^3: one timed step "
                  f"after one warm-up ({dt1:.1f} s, {rate1:.3e} voxel-steps/s at {n}^3), scaled to {n_target}^3 by "
                  f"N log N; 1 thread of {os.cpu_count()} on '{cpu_model()}'",
        "measured_rate_at_sample": rate1,
        "multicore": {"value": ratem * (n_target ** 3 / n ** 3) / scale, "cores": ncores, "measured_rate_at_sample": ratem,
                      "sample": f"same, scipy.fft with {ncores} workers ({dtm:.1f} s per step at {n}^3); "
                                f"element-wise NumPy work stays single-threaded"},
    }


def scan_benchmark(count=96, seed=4):
    """Uneven 2-D sweep: `count` scan-line simulations batched vs run one at a time on the same GPU."""
    import numpy as np

    from kwave_amd import workloads
    from kwave_amd.batch import BatchSimulation
    from kwave_amd.solvers.kspace_solver import collect_results
    from kwave_amd.solvers.plan import DevicePlan

    t0 = time.perf_counter()
    cases = workloads.scan_sweep_2d(seed=seed, count=count)
    bs = BatchSimulation(cases)
    log(f"scan: {count} instances in {len(bs.plans)} size groups, set up in {time.perf_counter() - t0:.1f} s")
    voxel_steps = sum(int(np.prod(s.shape)) * s.Nt for s in bs.setups)
    t0 = time.perf_counter()
    batched = bs.run()
    t_batched = time.perf_counter() - t0
    # per size group: grid, instances, the column tile widths and radix plans the plan's timing picked
    groups = [{"grid": list(bs.setups[idx[0]].shape), "instances": len(idx), "tiles": list(pl.tiles()),
               "radix_plans": list(pl.radix_plans())} for pl, idx in zip(bs.plans, bs.groups)]
    bs.close()
    log(f"scan: batched {t_batched:.2f} s")
    singles = [DevicePlan([st]) for st in bs.setups]
    t0 = time.perf_counter()
    for pl, st in zip(singles, bs.setups):
        pl.run(st.Nt - pl.step_index(0))
    t_looped = time.perf_counter() - t0
    diff = 0.0
    for pl, st, res in zip(singles, bs.setups, batched):
        one = collect_results(pl, 0, st, np.float64)
        diff = max(diff, float(np.max(np.abs(one["p"] - res["p"])) / max(np.max(np.abs(one["p"])), 1e-30)))
        pl.close()
    log(f"scan: looped {t_looped:.2f} s, speedup {t_looped / t_batched:.2f}x, max rel diff {diff:.1e}")
    return {"instances": count, "size_groups": len(bs.groups), "groups": groups, "voxel_steps": voxel_steps,
            "batched_s": t_batched, "looped_s": t_looped, "speedup_vs_looped": t_looped / t_batched,
            "sims_per_s": count / t_batched, "voxel_steps_per_s": voxel_steps / t_batched,
            "max_rel_diff_batched_vs_looped": diff}


def small_workloads():
    """Throughput of the two 2-D single-instance workloads (128^2 IVP, 512^2 heterogeneous absorbing)."""
    import numpy as np

    from kwave_amd import workloads
    from kwave_amd.solvers.plan import DevicePlan
    from kwave_amd.solvers.setup import build_instance

    out = {}
    for name, case in (("2d_ivp_128", workloads.ivp_2d(seed=1)), ("2d_hetero_absorbing_512", workloads.hetero_2d(seed=2))):
        kg, med, src, sen, pml, alpha = case
        st = build_instance(kg, med, src, sen, pml_size=pml, pml_alpha=alpha)
        plan = DevicePlan([st])
        plan.run(st.Nt - plan.step_index(0))  # warm-up (graph capture)
        plan.reset()
        t0 = time.perf_counter()
        plan.run(st.Nt - plan.step_index(0))
        wall = time.perf_counter() - t0
        plan.close()
        vs = int(np.prod(st.shape)) * st.Nt
        out[name] = {"grid": list(st.shape), "nt": st.Nt, "wall_s": wall, "ms_per_step": 1e3 * wall / st.Nt,
                     "voxel_steps_per_s": vs / wall}
        log(f"{name}: {st.Nt} steps in {wall:.3f} s ({vs / wall:.3e} voxel-steps/s)")
    return out


def surrogate_leg():
    """Learned surrogate of the 128^2 problem (4 layers, 32 channels, 16 x 16 modes): train on 20k
    solver samples, calibrate the per-map fallback on 3k more, held-out error on 1k before and
    after routing, speed vs the batched solver with and without the fallback."""
    from kwave_amd import surrogate

    res, _ = surrogate.benchmark(n_train=20000, n_test=1000, n_cal=3000, epochs=36, log=log)
    return res


def adjoint_leg(checkpoint=50):
    """Reverse-time gradient dJ/dc0 on the 128^2, 600-step problem vs one forward run (same plan)."""
    from kwave_amd import _abi, workloads
    from kwave_amd.solvers.plan import DevicePlan
    from kwave_amd.solvers.setup import build_instance

    kg, med, src, sen, pml, alpha = workloads.surrogate_2d(seed=5_000_321)
    st = build_instance(kg, med, src, sen, pml_size=pml, pml_alpha=alpha)
    plan = DevicePlan([st])
    try:
        plan.run(st.Nt)
        plan.reset()
        plan.run(0)
        t0 = time.perf_counter()
        plan.run(st.Nt)
        t_fwd = time.perf_counter() - t0
        adj = 0.3 * plan.recorded(0, _abi.F_REC_P)
        plan.adjoint_c([adj], checkpoint=checkpoint)
        t0 = time.perf_counter()
        plan.adjoint_c([adj], checkpoint=checkpoint)
        t_adj = time.perf_counter() - t0
    finally:
        plan.close()
    log(f"adjoint: forward {1e3 * t_fwd:.1f} ms, gradient {1e3 * t_adj:.1f} ms ({t_adj / t_fwd:.1f}x)")
    return {"grid": list(st.shape), "nt": st.Nt, "checkpoint": checkpoint, "forward_ms": 1e3 * t_fwd,
            "gradient_ms": 1e3 * t_adj, "gradient_over_forward": t_adj / t_fwd}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=512, help="grid edge (512 is the benchmark workload)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-scan", action="store_true", help="skip the batched 96-instance 2-D sweep")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned-surrogate and adjoint legs")
    args = ap.parse_args()
    if args.gpus != 1:
        log("this benchmark runs on one GPU only: the 512^3 workload is a single simulation on one MI355X; "
            "multi-GPU sharding (slab decomposition with all-to-all transposes) is a design note, not a build")
        sys.exit(2)

    if not args.no_surrogate:
        import torch  # the surrogate leg uses PyTorch: initialise its GPU context before the solver library

        torch.cuda.init()
    from kwave_amd import workloads
    from kwave_amd.solvers.plan import DevicePlan
    from kwave_amd.solvers.setup import build_instance

    n = args.n
    nt_cfg = 1500
    steps = min(args.steps, nt_cfg)
    t0 = time.perf_counter()
    kg, med, src, sen, pml, alpha = workloads.nonlinear_3d(seed=3, n=n, nt=nt_cfg, pml=16 * n // 512 or 2)
    st = build_instance(kg, med, src, sen, pml_size=pml, pml_alpha=alpha)
    del med
    log(f"workload built in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    plan = DevicePlan([st])
    log(f"device plan ready in {time.perf_counter() - t0:.1f} s")

    if args.warmup:
        plan.run(min(args.warmup, nt_cfg))
        plan.reset()
    import ctypes

    from kwave_amd import _abi

    _abi.check(_abi.lib().kw_plan_run(plan._h, 0))  # synchronise
    t0 = time.perf_counter()
    kern = plan.run_profiled(steps)
    wall = time.perf_counter() - t0
    log(f"timed {steps} steps in {wall:.3f} s")
    ntot = n ** 3
    rate = ntot * steps / wall
    # dominant kernel
    dom = max(kern, key=lambda k: kern[k][0])
    ms, by, nl = kern[dom]
    avg_ms = ms / nl
    achieved = by / nl / (avg_ms * 1e-3) / 1e9
    kernels = {k: {"launches": v[2], "avg_us": 1e3 * v[0] / v[2], "total_ms": v[0],
                   "GBps": (v[1] / v[2]) / (v[0] / v[2] * 1e-3) / 1e9 if v[1] else None,
                   "bytes_per_launch": v[1] / v[2]} for k, v in kern.items()}
    alg_bytes_step = sum(v[1] for v in kern.values()) / steps
    gpu_ms_step = sum(v[0] for v in kern.values()) / steps
    tiles = plan.tiles()
    p_rec = plan.read(0, _abi.F_REC_P)
    finite = bool(abs(p_rec).max() < 1e12)
    out = {
        "metric": "voxel-steps/s, %HBM peak (512\u00b3); batched sims/s; surrogate speedup @2% err",
        "value": rate,
        "unit": "voxel-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"3d_nonlinear_{n}", "grid": [n, n, n], "nt": nt_cfg, "pml": pml[0],
                   "physics": "heterogeneous c/rho, B/A=6, power-law 0.5 dB/(MHz^1.1 cm)",
                   "source": "16x16 planar additive", "sensor": "256x256 plane, p + p_max"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom, n),
                     "alg_bytes_per_launch": by / nl},
        "hbm_frac_step_model": STEP_MODEL_BYTES * rate / (HBM_PEAK_GBS * 1e9),
        "alg_bytes_per_voxel_step": alg_bytes_step / ntot,
        # the step's own algorithmic bytes (every kernel's count, summed) over the step time
        "step_achieved_GBps": alg_bytes_step / ntot * rate / 1e9,
        "hbm_frac_step_algorithmic": alg_bytes_step / ntot * rate / (HBM_PEAK_GBS * 1e9),
        "gpu_ms_per_step": gpu_ms_step,
        "kernels": kernels,
        "column_tiles": dict(zip(("fft_only", "spectral_jobs"), tiles)),
        "finite": finite,
    }
    plan.close()
    if not args.no_scan:
        try:
            out["small_2d"] = small_workloads()
        except Exception as e:
            log(f"2-D workloads failed: {e!r}")
        try:
            out["batched_scan"] = scan_benchmark()
        except Exception as e:
            log(f"scan benchmark failed: {e!r}")
            out["batched_scan"] = None
    if not args.no_surrogate:
        for key, leg in (("surrogate", surrogate_leg), ("adjoint", adjoint_leg)):
            try:
                out[key] = leg()
            except Exception as e:
                log(f"{key} leg failed: {e!r}")
                out[key] = None
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline()
        except Exception as e:  # the GPU figure stands without it
            log(f"cpu baseline failed: {e!r}")
            out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
