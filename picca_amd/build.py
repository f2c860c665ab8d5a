"""Builds the engine's shared library ``picca_amd/libpicca_hip.so`` with hipcc for gfx950.

Usage: ``python -m picca_amd.build``.  The library is built in-tree so that it travels
with the source tree; nothing is installed.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBRARY = os.path.join(HERE, "libpicca_hip.so")
SOURCES = ["api.hip", "cont_api.hip", "mlp_api.hip", "metal_api.hip", "cov_api.hip",
           "wick_api.hip", "xi_kernel.hip", "xcf_kernel.hip", "dmat_kernel.hip",
           "metal_kernel.hip", "cov_kernel.hip", "wick_kernel.hip",
           "cont_kernel.hip", "mlp_kernel.hip",
           "neighbours.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # bin assignment must round like the reference's FP64 code: no FMA contraction
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    path = os.path.join(rocm, "bin", "hipcc")
    return path if os.path.exists(path) else "hipcc"


def build(verbose=False):
    sources = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIBRARY + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp] + sources + ["-lpthread"]
    if verbose:
        print(" ".join(cmd))
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError("hipcc failed building %s" % LIBRARY)
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True))
