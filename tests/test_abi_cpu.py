"""The C-ABI library loads and exports every entry point include/kwgpu.h declares (no GPU calls)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    text = (ROOT / "include" / "kwgpu.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(kw_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_boundary():
    names = _declared()
    for must in ("kw_plan_create", "kw_plan_set_instance", "kw_plan_set_field", "kw_plan_set_source", "kw_plan_set_sensor",
                 "kw_plan_init", "kw_plan_run", "kw_plan_read", "kw_plan_destroy", "kw_last_error", "kw_abi_version"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from kwave_amd import _abi

    lib = ctypes.CDLL(str(_abi.LIB_PATH))
    for name in _declared():
        assert hasattr(lib, name), name
    assert set(_declared()) <= set(_abi.EXPORTS) | {"kw_kernel_kind_count", "kw_kernel_kind_name"}


def test_abi_version_and_size_table():
    from kwave_amd import _abi

    L = _abi.lib()
    assert L.kw_abi_version() == 1
    # fused lengths, and every other length 2..4096 through the generic transforms
    for n in (64, 128, 256, 384, 512, 640, 768, 1024, 48, 40, 56, 100, 120, 144, 200, 243, 336, 1000, 2048, 13, 97, 4096, 2):
        assert _abi.size_supported(n), n
    for n in (1, 0, -3, 4097, 8192):
        assert not _abi.size_supported(n), n


def test_plan_creation_reports_bad_arguments_as_valueerror():
    """Argument validation happens before any device call."""
    from kwave_amd import _abi

    d = _abi.PlanDesc()
    d.ndim = 2
    d.n[0], d.n[1], d.n[2] = 5000, 64, 1
    d.spacing[0] = d.spacing[1] = d.spacing[2] = 1e-4
    d.batch = 1
    h = ctypes.c_void_p()
    with pytest.raises(ValueError, match="no transform"):
        _abi.check(_abi.lib().kw_plan_create(ctypes.byref(d), ctypes.byref(h)))
    d.n[0] = 64
    d.batch = 0
    with pytest.raises(ValueError, match="batch"):
        _abi.check(_abi.lib().kw_plan_create(ctypes.byref(d), ctypes.byref(h)))
