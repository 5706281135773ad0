"""Build the HIP library in-tree: goodphotons_amd/libgoodphotons.so (gfx950).

Every source is compiled by hipcc for gfx950 with packed-FP32 instructions
disabled, then linked into one shared library. Run `python -m
goodphotons_amd.build` or call build().
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJDIR = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libgoodphotons.so")

SOURCES = ["gp_kernels.hip", "gp_image.hip", "gp_bvh_device.hip", "gp_runtime.cpp", "gp_api.cpp", "gp_server.cpp", "gp_bvh.cpp", "gp_checkpoint.cpp", "gp_mirror.cpp"]
ARCH = "gfx950"


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (ROCm required to build the HIP library)")


def _flags(src: str) -> list[str]:
    f = ["-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-result", "-Wno-unused-function", "-Wno-unused-value",
         "-I", os.path.join(ROOT, "include"), "-I", CSRC]
    if src.endswith(".hip"):
        f += ["-x", "hip"]
    return f


def build(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    cc = hipcc()

    def compile_one(src):
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        cmd = [cc, *_flags(src), "-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        if verbose and r.stderr.strip():
            sys.stderr.write(r.stderr)
        return obj

    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_demo()
    return LIB


DEMO = os.path.join(PKG, "forward_demo")


def build_demo() -> str:
    """The C++ caller of host/gp_forward.hpp, linked against the library."""
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, "-std=c++20", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(PKG, "host"), os.path.join(PKG, "host", "forward_demo.cpp"),
           "-o", DEMO + ".tmp", "-L", PKG, "-l:libgoodphotons.so", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"demo build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(DEMO + ".tmp", DEMO)
    return DEMO


if __name__ == "__main__":
    print(build(verbose=True))
