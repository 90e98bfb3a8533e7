"""In-tree build of libropephysics.so for gfx950 with hipcc.

Numerics flags: -ffp-contract=off keeps a*b+c as two roundings, matching the
reference's CPU arithmetic; HIP's default IEEE division and square root are
kept (no fast-math). Packed FP32 instructions are disabled.

Each source compiles to its own object in parallel (ropephysics_amd/_build/),
then one link makes the library.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libropephysics.so")
SOURCES = ["rp_kernels.hip", "rp_kernels64.hip", "rp_host.cpp", "rp_batch.cpp", "rp_server.cpp"]
ARCH = os.environ.get("RP_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def _run(cmd):
    return subprocess.run(cmd, capture_output=True, text=True)


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ, exist_ok=True)
    objs = [os.path.join(OBJ, s + ".o") for s in SOURCES]
    cmds = [[hipcc] + FLAGS + ["-c", os.path.join(CSRC, s), "-o", o] for s, o in zip(SOURCES, objs)]
    with ThreadPoolExecutor(len(cmds)) as ex:
        results = list(ex.map(_run, cmds))
    tmp = OUT + ".tmp"
    if all(r.returncode == 0 for r in results):
        results.append(_run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC"] + objs + ["-o", tmp]))
    for r in results:
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("hipcc failed building libropephysics.so")
        if verbose:
            sys.stderr.write(r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
