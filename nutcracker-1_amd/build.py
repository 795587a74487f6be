"""Build libnutcracker_hip.so for gfx950 with hipcc (no CMake, no JIT cache).

    python nutcracker-1_amd/build.py [--jobs N] [--verbose]

Object files go to nutcracker-1_amd/_build/, the shared library to
nutcracker-1_amd/nutcracker_amd/libnutcracker_hip.so (in-tree, so it travels with
the repository snapshot to the GPU host).
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shlex
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "nutcracker_amd", "libnutcracker_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

SOURCES = [
    "runtime.cpp",
    "kernels.hip",
    "gemm.hip",
    "planes.hip",
    "contract.cpp",
    "eigen_host.cpp",
    "iram.cpp",
    "optimize.cpp",
    "polar.cpp",
    "overlap.cpp",
    "chain.cpp",
    "batch.cpp",
    "abi.cpp",
    "abi_support.cpp",
    "abi_noalloc.cpp",
    "residency.cpp",
    "chain_abi.cpp",
    "bench_abi.cpp",
    "formats.cpp",
    "c_interface.cpp",
]

CXXFLAGS = [
    "-x", "hip",
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    # packed FP32 instructions off (they have produced silent wrong results on
    # these machines)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(HERE, "..", "include"),
]
LDFLAGS = [
    "-shared",
    "--offload-arch=gfx950",
    "-L", os.path.join(ROCM, "lib"),
    "-lrocsolver",
    "-lrocblas",
    "-Wl,-rpath," + os.path.join(ROCM, "lib"),
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(shlex.quote(c) for c in cmd), flush=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout)
        raise RuntimeError("command failed: " + " ".join(cmd[:3]) + " ...")
    return p.stdout


def build(jobs: int = 8, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC] + CXXFLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for out in ex.map(lambda c: _run(c, verbose), cmds):
            if out.strip() and verbose:
                print(out)
    tmp = OUT + ".tmp"
    _run([HIPCC] + objs + LDFLAGS + ["-o", tmp], verbose)
    os.replace(tmp, OUT)
    return OUT


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))


if __name__ == "__main__":
    main()
