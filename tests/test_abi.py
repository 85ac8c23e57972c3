"""C-ABI checks that need no GPU: the library loads, exports every declared symbol, and
rejects bad input with the documented status codes before touching a device."""
import ctypes

import pytest

from gwhip import lib as gl


def test_library_exports_every_header_symbol():
    lib = gl.load()
    names = gl.header_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n


def test_reader_library_exports_every_header_symbol():
    h5 = gl.load_h5()
    names = gl.header_functions(gl.H5_HEADER)
    assert len(names) >= 7
    for n in names:
        assert hasattr(h5, n), n


def test_debug_library_exports_every_header_symbol():
    """The probes live in their own library; libgwhip.so exports none of them."""
    dbg = gl.load_debug()
    names = gl.header_functions(gl.DEBUG_HEADER)
    assert names == ["gwhip_debug_gemm_bench", "gwhip_debug_mfma4_layout", "gwhip_debug_mfma_peak"]
    lib = gl.load()
    for n in names:
        assert hasattr(dbg, n), n
        assert not hasattr(lib, n), n


def test_headers_are_plain_c():
    """The boundary headers compile as C99 and as C++17 (a C caller and the C++ adapter both
    include them), with warnings as errors."""
    import os
    import shutil
    import subprocess
    import tempfile

    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    src = '#include "gwhip.h"\n#include "gwhip_h5.h"\n#include "gwhip_debug.h"\nint main(void) { return GWHIP_OK; }\n'
    with tempfile.TemporaryDirectory() as d:
        for compiler, ext, std in (("gcc", "c", "-std=c99"), ("g++", "cpp", "-std=c++17")):
            if not shutil.which(compiler):
                pytest.skip(f"{compiler} not found")
            path = os.path.join(d, "probe." + ext)
            with open(path, "w") as f:
                f.write(src)
            r = subprocess.run([compiler, std, "-Wall", "-Wextra", "-pedantic", "-Werror", "-fsyntax-only", "-I", inc,
                                path], capture_output=True, text=True)
            assert r.returncode == 0, r.stderr


def _create(**kw):
    lib = gl.load()
    base = dict(nk=1, mesh=(1, 1, 1), nao=2, nso=2, ns=1, NQ=3, nts=4, nw_b=3, beta=10.0)
    base.update(kw)
    d = gl.Dims(base["nk"], (ctypes.c_int64 * 3)(*base["mesh"]), base["nao"], base["nso"], base["ns"],
                base["NQ"], base["nts"], base["nw_b"], base["beta"])
    h = ctypes.c_void_p()
    st = lib.gwhip_create(ctypes.byref(d), 0, ctypes.byref(h))
    if st == gl.OK:
        lib.gwhip_destroy(h)
    return st, lib.gwhip_last_error(None).decode()


def test_odd_nts_is_wrong_grid():
    st, msg = _create(nts=5)
    assert st == gl.ERR_WRONG_GRID and "even" in msg


@pytest.mark.parametrize("bad", [dict(nk=0), dict(mesh=(2, 1, 1)), dict(nso=3), dict(ns=3),
                                 dict(ns=2, nso=4), dict(NQ=-1)])
def test_bad_dims_are_rejected(bad):
    st, _ = _create(**bad)
    assert st == gl.ERR_ARG


def test_null_arguments():
    lib = gl.load()
    assert lib.gwhip_create(None, 0, None) == gl.ERR_ARG
    assert lib.gwhip_set_grid(None, None, None) == gl.ERR_ARG
    assert lib.gwhip_eval_sigma(None, None, None, 0) == gl.ERR_ARG
    assert lib.gwhip_device_bytes(None) == 0


def test_valid_dims_reach_the_device_check():
    """Without a GPU the valid request fails only at the device step (no silent fallback)."""
    st, _ = _create()
    assert st in (gl.OK, gl.ERR_DEVICE)
