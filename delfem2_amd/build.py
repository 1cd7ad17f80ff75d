"""Builds the gfx950 shared library ``delfem2_amd/libdfm2_hip.so`` in-tree.

One hipcc invocation over every ``csrc/*.hip``.  Flags:
* ``--offload-arch=gfx950`` -- CDNA4 only, no other targets;
* ``-ffp-contract=off`` -- no fused multiply-add anywhere, so the FP64 kernels
  round exactly like the (non-FMA) CPU reference;
* ``-Xclang -target-feature -Xclang -packed-fp32-ops`` -- keeps the compiler
  from emitting packed FP32 instructions.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libdfm2_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def sources():
    return sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))


def build(verbose: bool = False, out: str = LIB, defines=()) -> str:
    srcs = sources()
    tmp = out + ".tmp"
    cmd = [HIPCC, *FLAGS, *[f"-D{d}" for d in defines], "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libdfm2_hip.so")
    # warnings from the host pass about the packed-fp32 feature are expected
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
