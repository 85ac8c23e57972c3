"""Build libgwhip.so (HIP, gfx950) in-tree.

    python green-mbpt_amd/build.py            # compile
    python green-mbpt_amd/build.py --print    # show the command only

The library lands next to the Python bindings (green-mbpt_amd/gwhip/libgwhip.so) so it
travels with the source tree.  Packed-FP32 instructions are disabled for all device code.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "gwhip", "libgwhip.so")
SOURCES = ["gw_hip_kernel.hip"]
ARCH = os.environ.get("GWHIP_ARCH", "gfx950")


def command(out: str = OUT) -> list[str]:
    hipcc = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")
    return [hipcc, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(ROOT, "include"), "-I", CSRC,
            "-o", out] + [os.path.join(CSRC, s) for s in SOURCES]


def build(verbose: bool = True) -> str:
    cmd = command(OUT + ".tmp")
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    if "--print" in sys.argv:
        print(" ".join(command()))
    else:
        build()
