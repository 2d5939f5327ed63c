"""Build the HIP library ``libcbs_mi355x.so`` for gfx950 (in tree).

Compiles every ``csrc/*.hip`` unit with hipcc in parallel and links the
shared library next to this file (``lib/libcbs_mi355x.so``).  Objects go to
``frequencymaxwell.jl_amd/_build``.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = ROOT / "csrc"
BUILD = ROOT / "_build"
LIBDIR = PKG / "lib"
LIBNAME = "libcbs_mi355x.so"
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions stay off (the compiler would otherwise form
    # v_pk_* FP32 ops from scalar and float2 code)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def lib_path() -> Path:
    return LIBDIR / LIBNAME


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: Path, verbose: bool, defines=(), objdir: Path = BUILD) -> Path:
    obj = objdir / (src.stem + ".o")
    cmd = [_hipcc(), *HIPCC_FLAGS, *[f"-D{d}" for d in defines], "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False, jobs: int | None = None, defines=(), out: Path | None = None) -> Path:
    """Compile all HIP units and link the shared library; returns its path.

    ``defines`` (``NAME=VALUE`` strings) and ``out`` build a tuning variant
    elsewhere; the default builds the library the package loads.
    """
    objdir = BUILD if not defines else BUILD / ("v_" + "_".join(d.replace("=", "") for d in defines))
    objdir.mkdir(parents=True, exist_ok=True)
    LIBDIR.mkdir(parents=True, exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    if not srcs:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    jobs = jobs or min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)), 16)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose, defines, objdir), srcs))
    out = Path(out) if out else lib_path()
    out.parent.mkdir(parents=True, exist_ok=True)
    tmp = out.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser()
    ap.add_argument("-v", action="store_true")
    ap.add_argument("-D", dest="defines", action="append", default=[])
    ap.add_argument("--out")
    a = ap.parse_args()
    print(build(verbose=a.v, defines=a.defines, out=a.out))
