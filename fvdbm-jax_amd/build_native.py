"""Compile the HIP library in-tree: fvdbm-jax_amd/_build/libfvdbm.so (gfx950).

Flags:
  -ffp-contract=off   every multiply and add rounds on its own (the
                      reference's operation order, bitwise reproducible)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions
  -mcode-object-version=5
                      loadable by the HIP runtime bundled with PyTorch-ROCm,
                      which is the runtime the library binds to (see _lib.py)
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "libfvdbm.so")

SOURCES = [
    os.path.join(HERE, "csrc", "fvdbm_step.hip"),
    os.path.join(HERE, "csrc", "fvdbm_lattice.hip"),
    os.path.join(HERE, "csrc", "surrogate.hip"),
    os.path.join(HERE, "csrc", "layout.cpp"),
    os.path.join(HERE, "csrc", "nn_inputs.cpp"),
]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mcode-object-version=5",
    "-Wall", "-Wno-unused-result",
]


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    return os.path.join(rocm, "bin", "hipcc")


def build(verbose=False):
    """Compile every HIP source into libfvdbm.so (always recompiles)."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(s)]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"),
                                     "-I", os.path.join(HERE, "csrc"),
                                     "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd))
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stderr[-8000:]}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
