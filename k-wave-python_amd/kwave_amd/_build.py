"""Build libkwgpu.so (gfx950) in-tree with hipcc.

Run as ``python -m kwave_amd._build`` (with k-wave-python_amd on sys.path) or
through ``__graft_entry__.build()``.  Objects are compiled in parallel and the
shared library lands in ``kwave_amd/_lib/libkwgpu.so``.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG.parent / "csrc"
INCLUDE = PKG.parent.parent / "include"
OUT_DIR = PKG / "_lib"
LIB = OUT_DIR / "libkwgpu.so"
# launcher files compiled once per line-length family (-DKW_GROUP=g), heaviest first
GROUPED = ["kw_zpass_mass.hip", "kw_col_pass.hip", "kw_zpass_mom.hip", "kw_zpass_eos.hip", "kw_zpass_io.hip", "kw_adjoint.hip"]
GROUPS = (3, 2, 1, 0)
SOURCES = ["kw_generic.hip", "kw_dispatch.hip", "kw_misc.hip", "kw_abi.hip", "kw_fno.hip", "kw_static.hip"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("KW_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    f"-I{CSRC}",
    f"-I{INCLUDE}",
]


def _compile(src: Path, obj: Path, defines=()):
    cmd = [HIPCC, *FLAGS, *defines, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    """Compile every source and link the shared library; returns its path."""
    build_dir = PKG.parent / "_build"
    build_dir.mkdir(exist_ok=True)
    OUT_DIR.mkdir(exist_ok=True)
    objs = []
    jobs = [(CSRC / s, build_dir / f"{Path(s).stem}.g{g}.o", (f"-DKW_GROUP={g}",)) for g in GROUPS for s in GROUPED]
    jobs += [(CSRC / s, build_dir / (Path(s).stem + ".o"), ()) for s in SOURCES]
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        futs = [ex.submit(_compile, *j) for j in jobs]
        for f in futs:
            objs.append(f.result())
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
