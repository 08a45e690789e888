"""Builds the native libraries in-tree.

  ish_amd/lib/libish_hip.so       C ABI + HIP kernels (hipcc, gfx950)
  ish_amd/lib/libish_workload.so  synthetic workload generator (host C++)

Run `python -m ish_amd.build`. Every call compiles again.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    # no packed fp32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

# the kernel files that take longest to compile come first
HIP_SOURCES = ["sg_hybrid.hip", "sg_multi.hip", "sg_lock.hip", "sg_batched.hip", "sw_kernels.hip",
               "sg_misc.hip", "ish_hip.cpp", "server.cpp", "search.cpp", "serve_kernels.hip",
               "noalloc.hip"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = True) -> None:
    os.makedirs(LIB, exist_ok=True)
    objs = [os.path.join(LIB, os.path.splitext(src)[0] + ".o") for src in HIP_SOURCES]
    # one compiler process per source file, at most 8 at a time
    cmds = [[hipcc(), *HIP_FLAGS(), "-c", os.path.join(CSRC, src), "-o", obj]
            for src, obj in zip(HIP_SOURCES, objs)]
    with ThreadPoolExecutor(max_workers=min(8, len(cmds))) as pool:
        for f in [pool.submit(_run, c) for c in cmds]:
            f.result()
    out = os.path.join(LIB, "libish_hip.so")
    tmp = out + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", *objs, "-lz", "-o", tmp])
    os.replace(tmp, out)
    wl = os.path.join(LIB, "libish_workload.so")
    _run(["g++", "-O3", "-fPIC", "-shared", "-std=c++17", "-pthread",
          os.path.join(CSRC, "workload.cpp"), "-o", wl + ".tmp"])
    os.replace(wl + ".tmp", wl)


def HIP_FLAGS():
    return list(HIPCC_FLAGS)


if __name__ == "__main__":
    build()
    sys.exit(0)
