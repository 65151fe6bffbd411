"""1-D problems as a 2 x N grid uniform along axis 0 (kwave_amd/solvers/embed1d.py): the host
side of the embedding (no GPU)."""
import numpy as np

from kwave_amd.kgrid import kWaveGrid
from kwave_amd.kmedium import kWaveMedium
from kwave_amd.ksensor import kSensor
from kwave_amd.ksource import kSource
from kwave_amd.solvers import embed1d
from kwave_amd.solvers.setup import build_instance


def _problem(per_point=False):
    kg = kWaveGrid([64], [1e-4])
    kg.setTime(50, 2e-8)
    c = 1500 + 50 * np.random.default_rng(1).random(64)
    src = kSource()
    m = np.zeros(64)
    m[[10, 12]] = 1
    src.p_mask = m
    base = np.sin(np.arange(20) * 0.3)
    src.p = np.stack([base, 2 * base]) if per_point else base.reshape(1, -1)
    sen = kSensor(np.zeros(64))
    sen.mask[[40, 50]] = 1
    sen.record = ["p", "u", "p_final", "ux_max", "I_avg"]
    return kg, kWaveMedium(sound_speed=c, density=1000.0), src, sen


def test_record_names_map_to_the_propagation_axis():
    assert embed1d.record_2d(["p", "u", "ux_staggered", "u_max", "ux_final", "I", "I_avg", "p_final"]) == [
        "p", "uy", "uy_staggered", "uy_max", "uy_final", "Iy", "Iy_avg", "p_final"]
    assert embed1d.record_2d(None) is None


def test_embedded_setup_puts_the_density_on_axis_1():
    kg, med, src, sen = _problem(per_point=True)
    e = embed1d.embed(kg, med, src, sen, (8,), (2.0,))
    st = embed1d.adjust_setup(build_instance(*e[:4], smooth_p0=False, pml_size=e[4], pml_alpha=e[5]))
    assert st.shape == (2, 64) and st.pml_sizes == [0, 8]
    assert np.array_equal(st.src_p.idx, [10, 12, 74, 76])          # both rows, C order
    assert st.src_p.sig.shape[0] == 4 and np.array_equal(st.src_p.sig[2:], st.src_p.sig[:2])
    # additive: 2 dt / (D c dx) with D = 2 in the embedding, so axis 1 takes twice its factor
    assert st.src_p.fac[0] == 0.0 and np.isclose(st.src_p.fac[1], 2 / 1e-4)
    assert np.allclose(st.src_p.scale * st.src_p.fac[1], 2 * 2e-8 / (np.asarray(med.sound_speed)[[10, 12, 10, 12]] * 1e-4))
    assert st.p0_split == (0.0, 1.0, 0.0)
    assert np.array_equal(st.sensor_idx.ravel(), [40, 50])         # row 0 only
    assert {"p", "uy", "p_final", "uy_max", "Iy_avg"} <= st.requested


def test_cartesian_sensor_sits_on_row_0():
    kg, med, src, _ = _problem()
    sen = kSensor(np.array([[1e-4 * 3.5, -1e-4 * 7.25]]))
    e = embed1d.embed(kg, med, src, sen, (8,), (2.0,))
    st = build_instance(*e[:4], smooth_p0=False, pml_size=e[4], pml_alpha=e[5])
    # corner pairs along axis 1 on row 0 with the 1-D linear weights; row 1 weights are 0
    w = st.sensor_w
    assert np.allclose(w.sum(axis=1), 1.0)
    row = st.sensor_idx // 64
    assert np.all(w[row == 1] == 0.0)
    assert np.allclose(np.sort(w[row == 0].reshape(2, 2), axis=1), [[0.5, 0.5], [0.25, 0.75]])


def test_unembed_reads_row_0_under_1d_names():
    res = {"p": np.ones((2, 5)), "uy": np.full((2, 5), 3.0), "ux": np.zeros((2, 5)),
           "p_final": np.arange(8.0).reshape(2, 4), "uy_final": np.arange(8.0).reshape(2, 4), "Iy_avg": np.ones(2)}
    out = embed1d.unembed(res)
    assert set(out) == {"p", "ux", "p_final", "ux_final", "Ix_avg"}
    assert out["ux"].shape == (2, 5) and np.all(out["ux"] == 3.0)        # series keep (n_sensor, Nt)
    assert np.array_equal(out["p_final"], [0, 1, 2, 3])
