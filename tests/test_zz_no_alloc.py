"""The allocation-free spectral operator (kw_spectral_apply_static) against kw_spectral_apply.

Both run the same kernels; the allocation-free entry point keeps its spectrum and
tables in a fixed per-GPU device array, so its results must equal the allocating
entry point's bit for bit.  Two calls in a row, the second on a smaller grid (the
array keeps the first call's contents), and a grid beyond the array's size is
refused before any device work.  Device arrays for the input and output come from
PyTorch (allocated before the calls; the entry point itself allocates nothing).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _static_apply(f, spacing, kind, axis=0, kap=0.0, use_sg=True, power=0.0):
    import torch

    from kwave_amd import _abi

    nd = f.ndim
    n = (C.c_int * 3)(*(list(f.shape) + [1] * (3 - nd)))
    d = (C.c_double * 3)(*(list(spacing) + [1.0] * (3 - nd)))
    fin = torch.from_numpy(np.ascontiguousarray(f, dtype=np.float32)).to("cuda")
    out = torch.full_like(fin, float("nan"))
    torch.cuda.synchronize()
    _abi.check(_abi.lib().kw_spectral_apply_static(nd, n, d, int(use_sg), float(kap), int(kind), int(axis), float(power),
                                                   C.c_void_p(fin.data_ptr()), C.c_void_p(out.data_ptr()), None))
    return out.cpu().numpy()


@pytest.mark.parametrize("order", ["large_then_small", "2d_then_3d"])
def test_static_entry_point_equals_allocating_one(order):
    from kwave_amd import _abi

    rng = np.random.default_rng(17)
    cases = [((64, 48, 96), (1e-4, 2e-4, 1.5e-4), 1, 1, 0.31, 0.0),   # gradient along y
             ((40, 56), (1e-4, 1e-4), 3, 0, 0.12, -0.9)]              # |k|^-0.9
    if order == "2d_then_3d":
        cases = [((32, 24), (5e-5, 5e-5), 2, 1, 0.2, 0.0), ((24, 32, 16), (1e-4, 1e-4, 1e-4), 5, 0, 0.4, 0.0)]
    for shape, sp, kind, axis, kap, power in cases:
        f = rng.standard_normal(shape).astype(np.float32)
        want = _abi.spectral_apply(f, sp, kind, axis=axis, kap=kap, power=power)
        got = _static_apply(f, sp, kind, axis=axis, kap=kap, power=power)
        print(shape, "kind", kind, "max |diff|", float(np.max(np.abs(got - want))))
        assert np.array_equal(got, want)


def test_static_entry_point_refuses_grids_beyond_its_scratch():
    import torch

    from kwave_amd import _abi

    n = (C.c_int * 3)(256, 256, 256)
    d = (C.c_double * 3)(1e-4, 1e-4, 1e-4)
    buf = torch.zeros(8, device="cuda")  # never touched: the size check comes first
    rc = _abi.lib().kw_spectral_apply_static(3, n, d, 1, 0.0, 1, 0, 0.0, C.c_void_p(buf.data_ptr()),
                                             C.c_void_p(buf.data_ptr()), None)
    assert rc == _abi.KW_EINVAL
    assert "scratch" in _abi.lib().kw_last_error().decode()
