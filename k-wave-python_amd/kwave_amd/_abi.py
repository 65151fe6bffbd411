"""ctypes binding of libkwgpu.so (declared in include/kwgpu.h).

The library is loaded from ``kwave_amd/_lib/libkwgpu.so``; there is no
fallback: a missing library raises ImportError, so the HIP path is the only
path.
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libkwgpu.so"

KW_OK, KW_EINVAL, KW_ENOMEM, KW_EHIP, KW_EUNSUPPORTED = 0, -1, -2, -3, -4
KW_NONLINEAR, KW_STOKES, KW_POWERLAW_ABS, KW_POWERLAW_DISP = 1, 2, 4, 8
F_C0SQ, F_RHO0, F_DTR0_X, F_BONA, F_TAU, F_ETA, F_P0 = 0, 1, 2, 5, 6, 7, 8
F_P, F_UX, F_RHOX, F_REC_P, F_REC_UX, F_AGG_P, F_STEP, F_ALPHA_NP, F_REC_UCX = 9, 10, 13, 16, 17, 20, 21, 22, 23
SRC_DIRICHLET, SRC_ADDITIVE, SRC_ADDITIVE_NOCORR = 1, 2, 4
REC_P, REC_P_AGG, REC_UX, REC_UCX = 1, 2, 4, 32

# every symbol include/kwgpu.h declares
EXPORTS = (
    "kw_abi_version", "kw_last_error", "kw_device_count", "kw_size_supported",
    "kw_plan_create", "kw_plan_set_instance", "kw_plan_set_field", "kw_plan_set_source",
    "kw_plan_set_sensor", "kw_plan_init", "kw_plan_run", "kw_plan_read", "kw_plan_field_size",
    "kw_plan_destroy", "kw_spectral_apply", "kw_plan_run_profiled", "kw_plan_launch", "kw_plan_sync", "kw_kernel_kind_count", "kw_kernel_kind_name",
    "kw_plan_adjoint_c", "kw_mode_gemm_bf16", "kw_mode_contract_bf16", "kw_plan_tiles", "kw_plan_radix_plans", "kw_plan_set_density_split", "kw_fno_pointwise_bf16",
    "kw_fno_dft_bf16", "kw_fno_head_gather_bf16", "kw_fno_lift_bf16", "kw_spectral_apply_static", "kw_batch_create", "kw_batch_run",
    "kw_batch_destroy",
)


class PlanDesc(C.Structure):
    _fields_ = [
        ("ndim", C.c_int), ("n", C.c_int * 3), ("spacing", C.c_double * 3), ("batch", C.c_int),
        ("flags", C.c_int), ("alpha_power", C.c_double), ("use_sg", C.c_int), ("device", C.c_int),
    ]


class InstDesc(C.Structure):
    _fields_ = [
        ("nt", C.c_int), ("dt", C.c_double), ("c_ref", C.c_double), ("use_kspace", C.c_int),
        ("pml", C.POINTER(C.c_float) * 3), ("pml_sg", C.POINTER(C.c_float) * 3),
    ]


_lib = None


def _torch_context_first():
    """Initialise PyTorch's GPU context first when PyTorch is in the process.

    The PyTorch wheel carries its own HIP runtime next to the system one this
    library links; with both in one process, PyTorch has found no GPU when this
    library initialised the device first, while the opposite order works.
    """
    torch = sys.modules.get("torch")
    if torch is None:
        return
    try:
        if torch.cuda.is_available():
            torch.cuda.init()
    except Exception:  # a CPU-only PyTorch: nothing to order
        pass


def lib():
    """Load (once) and return the shared library with typed signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m kwave_amd._build` (hipcc, gfx950)")
    _torch_context_first()
    L = C.CDLL(str(LIB_PATH))
    vp, i32, i64, f32p, f64p, i64p = C.c_void_p, C.c_int, C.c_int64, C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int64)
    sig = {
        "kw_abi_version": ([], i32),
        "kw_last_error": ([], C.c_char_p),
        "kw_device_count": ([C.POINTER(C.c_int)], i32),
        "kw_size_supported": ([i32], i32),
        "kw_plan_create": ([C.POINTER(PlanDesc), C.POINTER(vp)], i32),
        "kw_plan_set_instance": ([vp, i32, C.POINTER(InstDesc)], i32),
        "kw_plan_set_field": ([vp, i32, i32, f32p, i64], i32),
        "kw_plan_set_source": ([vp, i32, i32, i32, i64, i64p, f32p, C.POINTER(C.c_int32), f32p, i32, i32, f64p], i32),
        "kw_plan_set_sensor": ([vp, i32, i32, i32, i64p, f32p, i32, C.c_uint], i32),
        "kw_plan_init": ([vp], i32),
        "kw_plan_run": ([vp, i64], i32),
        "kw_plan_launch": ([vp, i64], i32),
        "kw_plan_sync": ([vp], i32),
        "kw_batch_create": ([C.POINTER(vp), i32, C.POINTER(vp)], i32),
        "kw_batch_run": ([vp], i32),
        "kw_batch_destroy": ([vp], i32),
        "kw_plan_read": ([vp, i32, i32, f32p, i64], i32),
        "kw_plan_field_size": ([vp, i32, i32, C.POINTER(C.c_int64)], i32),
        "kw_plan_destroy": ([vp], i32),
        "kw_spectral_apply": ([i32, C.POINTER(C.c_int), f64p, i32, C.c_double, i32, i32, C.c_double, f32p, f32p, i32], i32),
        "kw_spectral_apply_static": ([i32, C.POINTER(C.c_int), f64p, i32, C.c_double, i32, i32, C.c_double, vp, vp, vp], i32),
        "kw_plan_run_profiled": ([vp, i64, f64p, f64p, i64p], i32),
        "kw_plan_adjoint_c": ([vp, f32p, i32, f32p], i32),
        "kw_plan_tiles": ([vp, C.POINTER(C.c_int), C.POINTER(C.c_int)], i32),
        "kw_plan_radix_plans": ([vp, C.POINTER(C.c_int), C.POINTER(C.c_int)], i32),
        "kw_plan_set_density_split": ([vp, i32, f64p], i32),
        "kw_mode_gemm_bf16": ([i32, i32, i32, i32, vp, vp, vp, vp], i32),
        "kw_mode_contract_bf16": ([i32, i32, i32, i32, i32, vp, i64p, vp, vp, i64p, vp], i32),
        "kw_fno_pointwise_bf16": ([i32, i32, i64, vp, vp, vp, vp, vp, i32, vp], i32),
        "kw_fno_dft_bf16": ([i32, i64, vp, vp, vp, vp, vp], i32),
        "kw_fno_head_gather_bf16": ([i64, i32, i32, vp, vp, vp, vp], i32),
        "kw_fno_lift_bf16": ([i64, i32, i32, vp, C.c_float, C.c_float, vp, vp, vp, vp], i32),
        "kw_kernel_kind_count": ([], i32),
        "kw_kernel_kind_name": ([i32], C.c_char_p),
    }
    for name, (args, ret) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ret
    if L.kw_abi_version() != 1:
        raise ImportError("libkwgpu ABI version mismatch")
    _lib = L
    return L


def check(rc):
    """Map a KW_* return code to the Python exception the reference would raise."""
    if rc == KW_OK:
        return
    msg = lib().kw_last_error().decode(errors="replace")
    if rc == KW_ENOMEM:
        raise MemoryError(msg)
    if rc in (KW_EINVAL, KW_EUNSUPPORTED):
        raise ValueError(msg)
    raise RuntimeError(f"HIP error: {msg}")


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def device_count():
    n = C.c_int(0)
    check(lib().kw_device_count(C.byref(n)))
    return n.value


def size_supported(n):
    return bool(lib().kw_size_supported(int(n)))


def spectral_apply(f, spacing, kind, axis=0, kap=0.0, use_sg=True, power=0.0, device=0):
    """Device real(ifftn(op * fftn(f))) for one field (test hook for the transform kernels)."""
    f = f32(f)
    nd = f.ndim
    n = (C.c_int * 3)(*(list(f.shape) + [1] * (3 - nd)))
    d = (C.c_double * 3)(*(list(spacing) + [1.0] * (3 - nd)))
    out = np.empty_like(f)
    check(lib().kw_spectral_apply(nd, n, d, int(use_sg), float(kap), int(kind), int(axis), float(power), fptr(f), fptr(out), int(device)))
    return out
