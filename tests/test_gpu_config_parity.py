"""Per-configuration parity report: HIP solver vs the FP64 oracle on the frozen workloads.

Each case is one of the benchmark workloads of kwave_amd/workloads.py, at its
real size where the FP64 oracle finishes in seconds and shrunk otherwise (the
shrink is printed).  Relative L2 and max-abs/peak of every recorded output
are printed and the relative L2 must stay below 1e-4 (north-star bar for FP32
arithmetic against the FP64 reference solver).  The oracle restates
kwave/solvers/kspace_solver.py:122-862.

  2d_ivp_128      the 128^2 IVP at full size (664 steps, 50 Cartesian points)
  2d_hetero_*     the 512^2 absorbing line-source case shrunk to 128^2 (900 steps)
                  and 192^2 (1100 steps, a mixed-radix length), long enough for
                  the wave to cross from the source line to the sensor line
  3d_nonlinear    the 512^3 nonlinear absorbing case at 40^3, 300 steps (40 = 5 * 8;
                  the burst has reached the sensor plane and peaked)
  scan_<size>     one scan line of the 96-instance sweep per grid size, 150 steps
"""
import json
import os

import numpy as np
import pytest

from test_gpu_parity import REL_L2, _both, _compare

pytestmark = pytest.mark.gpu


CASES = ["2d_ivp_128", "2d_hetero_128_nt900", "2d_hetero_192_nt1100", "3d_nonlinear_40_nt300"] + [
    f"scan_{a}x{b}_nt150" for a, b in ((256, 256), (384, 512), (512, 768), (640, 640))]


def _case(name):
    from kwave_amd import workloads

    if name == "2d_ivp_128":
        return workloads.ivp_2d(seed=1)
    if name.startswith("2d_hetero_"):
        n, nt = (int(v) for v in name[len("2d_hetero_"):].split("_nt"))
        return workloads.hetero_2d(seed=2, n=n, nt=nt)
    if name == "3d_nonlinear_40_nt300":
        return workloads.nonlinear_3d(seed=3, n=40, nt=300, pml=3)
    a, b = (int(v) for v in name.split("_")[1].split("x"))
    return workloads.scan_line_2d(4000 + a, shape=(a, b), nt=150)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("name", CASES)
def test_config_parity(name):
    got, want = _both(_case(name))
    res = _compare(got, want, bar=REL_L2)
    out = os.environ.get("KW_PARITY_REPORT")
    if out:
        rep = json.load(open(out)) if os.path.exists(out) else {}
        rep[name] = {k: {"rel_l2": float(r), "max_abs_over_peak": float(m)} for k, (r, m) in res.items()}
        with open(out, "w") as f:
            json.dump(rep, f, indent=1)
