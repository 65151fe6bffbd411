"""Plane-wave physics on the HIP path itself (tests/physics_cases.py), with the bars of the
oracle pins (tests/test_oracle_physics.py): power-law decay and Kramers-Kronig phase speed
for y = 1.1, 1.5 and Stokes y = 2, second-harmonic growth with beta = 1 + B/2A, and the
plane-wave intensity.  The y-uniform plane wave runs on 16 grid rows here (4 for the
oracle: the field does not depend on the row count)."""
import copy

import numpy as np
import pytest

import physics_cases as pc
from test_oracle_physics import AMP_TOL, BETA_TOL, DISP_TOL, INT_TOL

pytestmark = pytest.mark.gpu


def _gpu(kg, med, src, sen, pml_size, pml_alpha):
    from kwave_amd.solvers.kspace_solver import Simulation

    sim = Simulation(kg, med, copy.deepcopy(src), sen, pml_size=pml_size, pml_alpha=pml_alpha, smooth_p0=False)
    try:
        return sim.run()
    finally:
        sim.close()


@pytest.mark.parametrize("y,alpha_db", [(1.1, 0.5), (1.5, 1.5), (2.0, 0.5)])
def test_gpu_plane_wave_decay_and_kramers_kronig_speed(y, alpha_db):
    amp_err, disp_err = pc.attenuation(_gpu, alpha_db, y, ny=16)
    print(f"y={y}: alpha error {amp_err.min():+.4f}..{amp_err.max():+.4f}, "
          f"dispersion error {disp_err.min():+.4f}..{disp_err.max():+.4f} (units of alpha)")
    assert np.all(np.abs(amp_err) < AMP_TOL)
    assert np.all(np.abs(disp_err) < DISP_TOL)


@pytest.mark.parametrize("bona", [6.0, 1e-4])
def test_gpu_second_harmonic_growth_gives_beta(bona):
    beta = pc.nonlinearity(_gpu, bona, ny=16)
    print(f"B/A={bona}: beta {beta:.4f} (expect {1 + bona / 2:.4f})")
    assert abs(beta / (1 + bona / 2) - 1) < BETA_TOL


def test_gpu_plane_wave_intensity():
    ratio, iy = pc.intensity(_gpu, ny=16)
    print(f"I_avg / (P^2 / 2 rho c) = {ratio:.5f}")
    assert abs(ratio - 1) < INT_TOL
    assert abs(iy) < 1e-5
