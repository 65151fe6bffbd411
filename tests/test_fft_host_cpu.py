"""Host check of the register FFT and its LDS index maps (no GPU).

tests/host_fft/fft_layouts.cpp compiles k-wave-python_amd/csrc/kw_fft.hpp as plain
C++ (tests/host_fft/hip/hip_runtime.h stands in for the HIP header: one POSIX
thread per lane, a barrier for the workgroup barrier) and runs fft_reg for every
compiled line length, both radix plans where a length has two, the column-pass
tile widths 1-32 (ColIdx) and the row-pass line counts 1-4 (RowIdx), in both
directions, against a direct double-precision DFT, with a guard band that flags
any exchange access outside the map's declared LDS size.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
CSRC = HERE.parent / "k-wave-python_amd" / "csrc"


@pytest.mark.timeout(600)
def test_register_fft_and_lds_maps_on_host(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = tmp_path / "fft_layouts"
    subprocess.run([gxx, "-std=c++17", "-O1", "-pthread", f"-I{HERE / 'host_fft'}", f"-I{CSRC}",
                    str(HERE / "host_fft" / "fft_layouts.cpp"), "-o", str(exe)], check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=500)
    bad = [ln for ln in r.stdout.splitlines() if ln.startswith("FAIL")]
    assert r.returncode == 0 and not bad, "\n".join(bad[:20]) or r.stdout[-2000:]
    assert r.stdout.strip().endswith("0 failed")
    assert sum(ln.startswith("ok") for ln in r.stdout.splitlines()) > 200
