"""Build libpefdtd.so (the gfx950 HIP kernels + C ABI) in-tree with hipcc.

Usage: python -m pulse_echo_azimuthal_simulation_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libpefdtd.so")
SOURCES = ["abi.hip", "kernels_direct.hip", "kernels_tiled.hip", "kernels_adjoint.hip",
           "kernels_adjoint_tiled.hip", "kernels_prep.hip"]
ARCH = os.environ.get("PEFDTD_ARCH", "gfx950")

# -ffp-contract=off: only the explicit fma() calls fuse, as in the oracle.
# -fgpu-flush-denormals-to-zero: f32 subnormals flushed (f64 keeps them),
#   the convention the upstream float32 archives were produced under.
# -packed-fp32-ops off: no v_pk_*_f32 instructions are generated.
HIPCC_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fgpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libpefdtd.so")


def build(verbose: bool = False) -> str:
    """Compile every HIP source into libpefdtd.so next to this file."""
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    cmd += [os.path.join(CSRC, s) for s in SOURCES] + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-q" not in sys.argv))
