"""Write FP64 oracle outputs for the GPU parity cases too slow to run beside the test.

    python tests/golden/make_oracle_fixtures.py [name ...]

Each fixture is an .npz of float32 (or float64) arrays next to this script:

* ``c3_box_{x,y,z}.npz``: the 512^3 workload's physics (B/A 6, power law
  y = 1.1, heterogeneous c/rho, additive plane source) on 512 x 32 x 32,
  32 x 512 x 32 and 32 x 32 x 512 (workloads.nonlinear_3d_box, seed 31, 300
  steps): p and p_max at every 4th point of the sensor plane.  These run the
  512-point kernels that the 512^3 benchmark times.
* ``c2_full.npz``: the 512^2 heterogeneous absorbing workload at full size
  (workloads.hetero_2d, seed 2, 2000 steps): p at every 4th sensor point.
* ``c5_fd.npz``: central finite differences (h = 0.1 m/s) of the trace misfit
  J(c) = 1/2 sum (p(c) - d)^2 of the 128^2, 600-step surrogate problem
  (workloads.surrogate_2d, seed 5_000_123) at 20 seeded cells (seed 7), with
  d = 0.7 p(c) + 5 % noise (seed 5), all from the FP64 oracle solver.

The oracle is test infrastructure; the fixtures are its outputs only.
"""
import copy
import sys
import time
from multiprocessing import Pool
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
for p in (ROOT, ROOT / "k-wave-python_amd"):
    sys.path.insert(0, str(p))

BOX_SHAPES = {"x": (512, 32, 32), "y": (32, 512, 32), "z": (32, 32, 512)}
SUB = 4  # keep every SUB-th sensor point


def _oracle_run(case):
    from oracle.kspace_ref import OracleSimulation

    kg, med, src, sen, pml, alpha = case
    return OracleSimulation(kg, med, copy.deepcopy(src), sen, pml_size=pml, pml_alpha=alpha).run()


def make_box(axis):
    from kwave_amd import workloads

    res = _oracle_run(workloads.nonlinear_3d_box(31, BOX_SHAPES[axis], nt=300))
    np.savez(HERE / f"c3_box_{axis}.npz", p=res["p"][::SUB].astype(np.float32),
             p_max=res["p_max"][::SUB].astype(np.float32), sub=SUB)


def make_c2():
    from kwave_amd import workloads

    res = _oracle_run(workloads.hetero_2d(seed=2))
    np.savez(HERE / "c2_full.npz", p=res["p"][::SUB].astype(np.float32), sub=SUB)


def c5_cells(c, seed=7, n=20):
    """20 cells in the region the wave reaches and returns from, away from the c maximum (c_ref fixed)."""
    rng = np.random.default_rng(seed)
    ii, jj = np.meshgrid(np.arange(c.shape[0]), np.arange(c.shape[1]), indexing="ij")
    near = (ii >= 22) & (ii < 64) & (jj >= 36) & (jj < 92) & (c < c.max() - 10.0)
    return rng.choice(np.where(near.ravel())[0], size=n, replace=False)


def _c5_J(args):
    from kwave_amd import workloads

    c, d = args
    res = _oracle_run(workloads.surrogate_2d(seed=5_000_123, c=c))
    return 0.5 * float(np.sum((res["p"] - d.astype(np.float64)) ** 2))


def make_c5(pool=None):
    from kwave_amd import workloads

    case = workloads.surrogate_2d(seed=5_000_123)
    c = np.asarray(case[1].sound_speed, dtype=np.float64)
    y = _oracle_run(case)["p"]
    d = (0.7 * y + 0.05 * np.abs(y).max() * np.random.default_rng(5).standard_normal(y.shape)).astype(np.float32)
    cells = c5_cells(c)
    h = 0.1
    jobs = []
    for cell in cells:
        for s in (+1, -1):
            cc = c.copy()
            cc.ravel()[cell] += s * h
            jobs.append((cc, d))
    Js = (pool.map(_c5_J, jobs) if pool else list(map(_c5_J, jobs)))
    fd = np.array([(Js[2 * i] - Js[2 * i + 1]) / (2 * h) for i in range(len(cells))])
    np.savez(HERE / "c5_fd.npz", cells=cells, fd=fd, d=d, h=h, J0=_c5_J((c, d)))


def main(names):
    t0 = time.time()
    with Pool(min(8, len(names) + 4)) as pool:
        waits = []
        for n in names:
            if n.startswith("c3_box_"):
                waits.append(pool.apply_async(make_box, (n[-1],)))
            elif n == "c2_full":
                waits.append(pool.apply_async(make_c2))
        if "c5_fd" in names:
            make_c5(pool)
        for w in waits:
            w.get()
    print(f"wrote {names} in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main(sys.argv[1:] or ["c3_box_x", "c3_box_y", "c3_box_z", "c2_full", "c5_fd"])
