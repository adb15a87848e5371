"""Build libbbpd_hip.so in-tree with hipcc for gfx950.

The library is compiled from bb_pd_amd/csrc/*.hip and *.cpp into bb_pd_amd/_build/.  Device code is
built with -ffp-contract=off (no fused multiply-add, so FP64 results follow the reference's
operation order bit for bit) and without packed FP32 instructions.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "_build"
LIB = OUT / "libbbpd_hip.so"
ARCH = os.environ.get("BBPD_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: cannot build libbbpd_hip.so")


def _compile(src: Path, obj: Path) -> None:
    cmd = [_hipcc(), *HIP_FLAGS, "-I", str(ROOT / "include"), "-I", str(CSRC), "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stdout}\n{r.stderr}")


def build_library(verbose: bool = False) -> Path:
    """Compile every source and link the shared library (always rebuilds)."""
    OUT.mkdir(exist_ok=True)
    srcs = sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))
    objs = [OUT / (s.stem + ".o") for s in srcs]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        list(ex.map(lambda p: _compile(*p), zip(srcs, objs)))
    tmp = OUT / "libbbpd_hip.so.tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


MEX = PKG / "mex"
MEX_OUT = OUT / "mex"
C_FLAGS = ["-std=c11", "-O2", "-fPIC", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__"]


def _rocm_include() -> str:
    return str(Path(_hipcc()).resolve().parent.parent / "include")


def _rocm_lib() -> str:
    return str(Path(_hipcc()).resolve().parent.parent / "lib")


def build_gateways(verbose: bool = False) -> list:
    """Compile the MEX gateways (bb_pd_amd/mex/*Mex.c) against the project's mex.h stand-in: one
    shared object per gateway, each exporting mexFunction, linked to libbbpd_hip.so and to
    libmexshim.so (the stand-in's implementation).  Needs the library built first."""
    MEX_OUT.mkdir(parents=True, exist_ok=True)
    inc = ["-I", str(ROOT / "include"), "-I", str(MEX), "-I", _rocm_include()]
    shim = MEX_OUT / "libmexshim.so"
    cmd = ["gcc", *C_FLAGS, *inc, "-shared", "-Wl,-soname,libmexshim.so", str(MEX / "mexshim.c"), "-o", str(shim)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gcc failed on mexshim.c:\n{r.stdout}\n{r.stderr}")
    out = []
    srcs = sorted(MEX.glob("*Mex.c"))

    def one(src):
        so = MEX_OUT / (src.stem + ".so")
        cmd = ["gcc", *C_FLAGS, *inc, "-shared", str(src), "-o", str(so), "-L", str(MEX_OUT), "-lmexshim",
               "-L", str(OUT), "-lbbpd_hip", "-L", _rocm_lib(), "-lamdhip64",
               "-Wl,-rpath,$ORIGIN:$ORIGIN/..:" + _rocm_lib()]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"gcc failed on {src.name}:\n{r.stdout}\n{r.stderr}")
        return so

    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        out = list(ex.map(one, srcs))
    if verbose:
        print(f"built {len(out)} MEX gateways in {MEX_OUT}")
    return out


if __name__ == "__main__":
    build_library(verbose=True)
    build_gateways(verbose=True)
