"""Batched uneven sweep: batch results equal the same instances run one at a time,
and one instance equals the FP64 oracle (rel-L2 <= 1e-4)."""
import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cases():
    from kwave_amd import workloads

    shapes = [(64, 96), (96, 96), (64, 96), (128, 80), (96, 96)]
    nts = [180, 240, 260, 200, 150]
    return [workloads.scan_line_2d(100 + i, shape=s, nt=nt, pml=10) for i, (s, nt) in enumerate(zip(shapes, nts))]


def test_batch_equals_looped_and_oracle():
    from kwave_amd.batch import BatchSimulation
    from kwave_amd.solvers.kspace_solver import Simulation
    from oracle.kspace_ref import OracleSimulation

    cases = _cases()
    bs = BatchSimulation(cases)
    assert len(bs.plans) == 3
    batched = bs.run()
    bs.close()
    for case, got in zip(cases, batched):
        kg, med, src, sen, pml, alpha = case
        one = Simulation(kg, med, copy.deepcopy(src), sen, pml_size=pml, pml_alpha=alpha).run()
        assert got["p"].shape == (32, kg.Nt)
        assert np.array_equal(got["p"], one["p"])
    kg, med, src, sen, pml, alpha = cases[3]
    want = OracleSimulation(kg, med, copy.deepcopy(src), sen, pml_size=pml, pml_alpha=alpha).run()
    rel = np.linalg.norm(batched[3]["p"] - want["p"]) / np.linalg.norm(want["p"])
    print("rel-L2 vs oracle", rel)
    assert rel < 1e-4


def test_batch_of_initial_value_problems():
    from kwave_amd import workloads
    from kwave_amd.batch import BatchSimulation
    from oracle.kspace_ref import OracleSimulation

    cases = [workloads.ivp_2d(seed=s, n=64, pml=10) for s in (21, 22, 23)]
    res = BatchSimulation(cases).run()
    for case, got in zip(cases, res):
        want = OracleSimulation(*case[:4], pml_size=case[4], pml_alpha=case[5]).run()
        rel = np.linalg.norm(got["p"] - want["p"]) / np.linalg.norm(want["p"])
        assert rel < 1e-4


def test_batch_with_different_alpha_power_equals_looped():
    """Absorbing instances that differ only in alpha_power land in different plans (the
    fractional-Laplacian exponents are per plan) and match single runs bit for bit."""
    from kwave_amd import workloads
    from kwave_amd.batch import BatchSimulation
    from kwave_amd.solvers.kspace_solver import Simulation

    cases = []
    for i, y in enumerate((1.1, 1.5, 1.1, 1.5)):
        kg, med, src, sen, pml, alpha = workloads.hetero_2d(seed=40 + i, n=96, nt=260, pml=10)
        med.alpha_power = y
        cases.append((kg, med, src, sen, pml, alpha))
    bs = BatchSimulation(cases)
    assert len(bs.plans) == 2
    batched = bs.run()
    bs.close()
    for case, got in zip(cases, batched):
        kg, med, src, sen, pml, alpha = case
        one = Simulation(kg, med, copy.deepcopy(src), sen, pml_size=pml, pml_alpha=alpha).run()
        assert np.array_equal(got["p"], one["p"])
    assert not np.allclose(batched[0]["p"], batched[1]["p"])


def test_batch_graph_equals_per_plan_graphs():
    """kw_batch_run (one graph per block of steps, size groups as parallel branches) and the
    plans replaying their own graphs give the same bits, with a second run on fresh plans."""
    from kwave_amd.batch import BatchSimulation

    cases = _cases()
    outs = []
    for graphs in ("batch", "plans", "batch"):
        bs = BatchSimulation(cases)
        outs.append(bs.run(graphs=graphs))
        bs.close()
    for a, b, c in zip(*outs):
        assert np.array_equal(a["p"], b["p"]) and np.array_equal(a["p"], c["p"])
