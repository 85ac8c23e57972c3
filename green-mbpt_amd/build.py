"""Build libgwhip.so (HIP, gfx950), libgwhip_debug.so (throughput probes, HIP) and libgwhip_h5.so
(host, HDF5 integral reader) in-tree.

    python green-mbpt_amd/build.py            # compile
    python green-mbpt_amd/build.py --print    # show the commands only

The libraries land next to the Python bindings (green-mbpt_amd/gwhip/) so they travel with the
source tree.  Packed-FP32 instructions are disabled for all device code.  libgwhip_h5.so links
the image's HDF5 (HDF5_ROOT, default /opt/conda) and libgwhip.so (found next to it).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "gwhip", "libgwhip.so")
OUT_H5 = os.path.join(HERE, "gwhip", "libgwhip_h5.so")
OUT_DEBUG = os.path.join(HERE, "gwhip", "libgwhip_debug.so")
HDF5_ROOT = os.environ.get("HDF5_ROOT", "/opt/conda")
SOURCES = ["gw_hip_kernel.hip", "gwhip_noalloc.hip"]
ARCH = os.environ.get("GWHIP_ARCH", "gfx950")


def command(out: str = OUT, sources=None) -> list[str]:
    hipcc = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")
    return [hipcc, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(ROOT, "include"), "-I", CSRC,
            "-o", out] + [os.path.join(CSRC, s) for s in (sources or SOURCES)]


def command_debug(out: str = OUT_DEBUG) -> list[str]:
    return command(out, ["gwhip_debug.hip"])


# libhdf5 is reached through a private directory holding only a link to it, so that the run path
# of libgwhip_h5.so cannot pull HDF5_ROOT's older libstdc++ in ahead of the system's.
H5_LINK_DIR = os.path.join(HERE, "gwhip", "hdf5")


def _hdf5_soname() -> str:
    lib = os.path.join(HDF5_ROOT, "lib")
    names = sorted(n for n in os.listdir(lib) if n.startswith("libhdf5.so.") and n.count(".") == 2)
    if not names:
        raise FileNotFoundError(f"no libhdf5.so.<N> under {lib}")
    return names[0]


def command_h5(out: str = OUT_H5) -> list[str]:
    return ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
            "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HDF5_ROOT, "include"),
            "-o", out, os.path.join(CSRC, "gw_h5_loader.cpp"),
            "-L", os.path.dirname(OUT), "-lgwhip", "-L", H5_LINK_DIR, "-l:" + _hdf5_soname(),
            "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/hdf5"]


def _link_hdf5():
    os.makedirs(H5_LINK_DIR, exist_ok=True)
    name = _hdf5_soname()
    link = os.path.join(H5_LINK_DIR, name)
    if os.path.islink(link) or os.path.exists(link):
        os.remove(link)
    os.symlink(os.path.join(HDF5_ROOT, "lib", name), link)


def build(verbose: bool = True) -> str:
    cmd = command(OUT + ".tmp")
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    cmd = command_debug(OUT_DEBUG + ".tmp")
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT_DEBUG + ".tmp", OUT_DEBUG)
    if os.path.exists(os.path.join(HDF5_ROOT, "include", "hdf5.h")):
        _link_hdf5()
        cmd = command_h5(OUT_H5 + ".tmp")
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(OUT_H5 + ".tmp", OUT_H5)
    elif verbose:
        print(f"no HDF5 under {HDF5_ROOT}: libgwhip_h5.so not built", flush=True)
    return OUT


if __name__ == "__main__":
    if "--print" in sys.argv:
        print(" ".join(command()))
        print(" ".join(command_debug()))
        print(" ".join(command_h5()))
    else:
        build()
