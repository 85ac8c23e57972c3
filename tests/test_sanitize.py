"""AddressSanitizer + UndefinedBehaviorSanitizer run of the host code at the boundary: the C ABI's
argument and state checks (the cases of test_abi.py, plus null handles on the other entry points)
and the HDF5 integral reader (gw_h5_loader.cpp) on the reference's integral and symmetry files.

The library is rebuilt with the sanitizers on its host code only (``-Xarch_host``; device code is
compiled as usual and never launched: without a GPU every call ends at its argument check or at the
device check), the reader and the driver (tests/sanitize/abi_h5_check.cpp) with the same clang, so
one sanitizer runtime serves the process.  Needs no GPU; takes about a minute (the library build).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "green-mbpt_amd", "csrc")
INC = os.path.join(ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HDF5_ROOT = os.environ.get("HDF5_ROOT", "/opt/conda")


def _hdf5_lib():
    lib = os.path.join(HDF5_ROOT, "lib")
    if not os.path.isdir(lib):
        return None
    names = sorted(n for n in os.listdir(lib) if n.startswith("libhdf5.so.") and n.count(".") == 2)
    return (lib, names[0]) if names else None


def test_boundary_under_asan_ubsan(tmp_path):
    hipcc, clang = os.path.join(ROCM, "bin", "hipcc"), os.path.join(ROCM, "llvm", "bin", "clang++")
    h5 = _hdf5_lib()
    if not (os.path.exists(hipcc) and os.path.exists(clang) and h5 and
            os.path.exists(os.path.join(HDF5_ROOT, "include", "hdf5.h"))):
        pytest.skip("hipcc, clang++ or HDF5 not found")
    d = str(tmp_path)
    # hdf5 through a directory holding only a link to it (keeps HDF5_ROOT's libstdc++ off the path)
    os.makedirs(os.path.join(d, "hdf5"))
    os.symlink(os.path.join(h5[0], h5[1]), os.path.join(d, "hdf5", h5[1]))
    lib = os.path.join(d, "libgwhip_san.so")
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined"]
    r = subprocess.run([hipcc, "--offload-arch=gfx950", "-O1", "-g", "-fno-omit-frame-pointer", "-std=c++17", "-fPIC",
                        "-shared"] + san + ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", INC,
                                            "-I", CSRC, "-o", lib, os.path.join(CSRC, "gw_hip_kernel.hip"),
                                            os.path.join(CSRC, "gwhip_noalloc.hip")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    exe = os.path.join(d, "abi_h5_check")
    r = subprocess.run([clang, "-O1", "-g", "-fno-omit-frame-pointer", "-std=c++17", "-fsanitize=address,undefined",
                        "-I", INC, "-I", os.path.join(HDF5_ROOT, "include"), "-o", exe,
                        os.path.join(ROOT, "tests", "sanitize", "abi_h5_check.cpp"), os.path.join(CSRC, "gw_h5_loader.cpp"),
                        "-L", d, "-lgwhip_san", "-L", os.path.join(d, "hdf5"), "-l:" + h5[1],
                        "-Wl,-rpath," + d, "-Wl,-rpath," + os.path.join(d, "hdf5")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               ASAN_OPTIONS="abort_on_error=0:halt_on_error=1")
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "h5")], capture_output=True, text=True, env=env,
                       timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-6000:]
    assert "all expectations hold" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    shutil.rmtree(d, ignore_errors=True)
