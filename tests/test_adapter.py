"""The green-mbpt adapter (green-mbpt_amd/adapter/green_gw_hip.h) compiles against declaration stubs
of the types it uses and behaves as the factory expects -- CPU only.

The stubs (tests/adapter_stubs/) stand in for green-mbpt's params, grids, symmetry, utils and
except headers.  factory_test.cpp calls the three hooks with the argument lists of
src/green/mbpt/kernel_factory.h:28-36 (custom_gw_kernel, custom_hf_kernel) and
src/green/mbpt/common_defs.h:169-180 (custom_kernel_parameters), checks that the reference's P_sp /
Sigma_sp flags select the precision (no flag of the kernel's own), and that without a GPU each hook
ends in mbpt_kernel_error rather than a silent CPU path.
"""
import os
import shutil
import subprocess

import pytest

from gwhip import lib as gl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STUBS = os.path.join(ROOT, "tests", "adapter_stubs")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def _compile(tmp_path, extra=()):
    if not shutil.which("g++"):
        pytest.skip("g++ not found")
    gl.load()
    gl.load_h5()  # both libraries must exist
    libdir = os.path.dirname(gl.LIB_PATH)
    exe = str(tmp_path / "factory_test")
    cmd = ["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", STUBS, "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "green-mbpt_amd", "adapter"), os.path.join(STUBS, "factory_test.cpp"),
           "-o", exe, "-L", libdir, "-lgwhip_h5", "-lgwhip", "-L", os.path.join(ROCM, "lib"),
           "-Wl,-rpath," + libdir, "-Wl,-rpath," + os.path.join(ROCM, "lib"), *extra]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_adapter_compiles_and_hooks_behave(tmp_path):
    exe = _compile(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout
    for line in ("precision switch of the kernel's own", "Sigma_sp=1 -> split Sigma",
                 "GW hook without a device throws", "HF solve without a device throws"):
        assert line in r.stdout
