"""Test infrastructure only: build the library's sources as host C++ against the CPU
emulation of HIP (hipemu.hpp), with AddressSanitizer, into tests/emu/_build/.

The kernel and runtime sources are compiled unchanged from goodphotons_amd/csrc; only
rocPRIM's device radix sort (used by the device BVH builder) is replaced by a stable
host sort (fakehip/rocprim). Used by tests/test_emulated_kernels.py to run small
traces with every memory access checked.
"""
from __future__ import annotations

import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "goodphotons_amd", "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libgp_emu.so")
CXX = "/opt/rocm/llvm/bin/clang++"
SOURCES = ["gp_kernels.hip", "gp_image.hip", "gp_bvh_device.hip", "gp_runtime.cpp", "gp_api.cpp",
           "gp_server.cpp", "gp_bvh.cpp", "gp_checkpoint.cpp", "gp_mirror.cpp"]


def asan_runtime() -> str:
    return subprocess.run([CXX, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True,
                          text=True, check=True).stdout.strip()


def available() -> bool:
    return os.path.exists(CXX) and os.path.exists(asan_runtime())


def build() -> str:
    os.makedirs(OUT, exist_ok=True)
    flags = ["-std=c++20", "-O2", "-g", "-fPIC", "-fno-omit-frame-pointer", "-fsanitize=address",
             "-I", os.path.join(HERE, "fakehip"), "-I", HERE, "-I", os.path.join(ROOT, "include"),
             "-I", CSRC, "-Wno-unused-result", "-Wno-ignored-attributes"]
    jobs = [(os.path.join(CSRC, s), os.path.join(OUT, os.path.splitext(s)[0] + ".o"), True) for s in SOURCES]
    jobs += [(os.path.join(HERE, "hipemu.cpp"), os.path.join(OUT, "hipemu.o"), False)]

    def one(job):
        src, obj, as_cxx = job
        cmd = [CXX, *flags, *(["-x", "c++"] if as_cxx else []), "-c", src, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"emulation build failed: {' '.join(cmd)}\n{r.stderr[-4000:]}")
        return obj

    with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(one, jobs))
    tmp = LIB + ".tmp"
    subprocess.run([CXX, "-shared", "-fsanitize=address", "-o", tmp, *objs, "-lpthread"], check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
