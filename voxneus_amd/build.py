"""Build libvoxneus_hip.so for gfx950 with hipcc (no CMake, no torch extension).

    python -m voxneus_amd.build [--jobs N] [--verbose]

Each ``csrc/*.hip`` file is compiled to an object in ``voxneus_amd/_build`` in parallel,
then linked into ``voxneus_amd/lib/libvoxneus_hip.so``.  Flags:

  * ``--offload-arch=gfx950``: MI355X (CDNA4) code only;
  * ``-ffp-contract=off``: every float expression rounds like the reference's PyTorch CPU
    ops (fused multiply-adds only where the kernels write ``fmaf`` explicitly);
  * ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed-FP32 instructions;
  * ``-munsafe-fp-atomics``: hardware float atomics on coarse-grained device memory.
"""

from __future__ import annotations

import argparse
import glob
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libvoxneus_hip.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VXN_OFFLOAD_ARCH", "gfx950")
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-fno-gpu-rdc", "-Wno-unused-result", f"-I{INCLUDE}",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{p.stdout}")
    if verbose and p.stdout.strip():
        print(p.stdout)
    return p.stdout


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def build(jobs: int = 0, verbose: bool = False) -> str:
    """Compile every HIP source for gfx950 and link the shared library; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    objs = [os.path.join(BUILD, os.path.basename(s) + ".o") for s in srcs]
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))

    def compile_one(pair):
        src, obj = pair
        _run([HIPCC, *CFLAGS, "-c", src, "-o", obj], verbose)
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(compile_one, zip(srcs, objs)))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    return LIB


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=0)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
    sys.exit(0)
