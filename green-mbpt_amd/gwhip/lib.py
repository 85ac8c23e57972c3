"""ctypes binding of libgwhip.so (include/gwhip.h) and a thin Python kernel object.

There is no CPU fallback: if the HIP library is missing or no GPU is visible, the calls
raise.  The object mirrors the reference's kernel life cycle
(``gw_kernel_factory::get_kernel`` -> kernel + callback, ``src/green/mbpt/kernel_factory.h:85-91``;
``solve(G, Sigma)``, ``src/gw_cpu_kernel.cpp:31-72``).
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgwhip.so")
H5_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgwhip_h5.so")
HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))),
                      "include", "gwhip.h")
H5_HEADER = os.path.join(os.path.dirname(HEADER), "gwhip_h5.h")
DEBUG_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgwhip_debug.so")
DEBUG_HEADER = os.path.join(os.path.dirname(HEADER), "gwhip_debug.h")
HDF5_ROOT = os.environ.get("HDF5_ROOT", "/opt/conda")

(OK, ERR_ARG, ERR_WRONG_GRID, ERR_STATE, ERR_OOM, ERR_DEVICE, ERR_NONFINITE, ERR_UNSUPPORTED, ERR_OUTDATED_INPUT,
 ERR_IO, ERR_MU_SEARCH, ERR_NOT_POSITIVE_DEFINITE) = range(12)
PREC_FP64, PREC_SPLIT, PREC_SPLIT_P0, PREC_SPLIT_SIGMA = 0, 1, 2, 3
VJP_RECOMPUTE, VJP_STORED = 0, 1
TUNING_KEYS = ("p0_tile", "sigma_tile", "aux_per_wave", "p0_group", "sigma_group", "p0_split_tile", "sigma_split_tile",
               "p0_fused_slices", "orbital_gemm", "p0_kchunks", "epi_form")
REGIONS = ("P0_zgemm", "P0(t) -> P0(w)", "GW-BSE", "P(w) -> P(t)", "Selfenergy_zgemm", "total")
KSTATS = ("P0 GEMM", "Sigma GEMM", "P0 prep", "Sigma epilogue", "Dyson inverse")


class GwhipError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"gwhip status {status}: {msg}")
        self.status = status


class WrongGrid(GwhipError):
    """Mirror of ``mbpt_wrong_grid`` (``src/green/mbpt/except.h:33-36``)."""


class OutdatedInput(GwhipError):
    """Mirror of ``mbpt_outdated_input`` (integral files older than format 1.0.0)."""


class ChemicalPotentialSearchFailure(GwhipError):
    """Mirror of ``mbpt_chemical_potential_search_failure``."""


def _raise(st: int, msg: str):
    if st == ERR_WRONG_GRID:
        raise WrongGrid(st, msg)
    if st == ERR_OUTDATED_INPUT:
        raise OutdatedInput(st, msg)
    if st == ERR_MU_SEARCH:
        raise ChemicalPotentialSearchFailure(st, msg)
    raise GwhipError(st, msg)


class Dims(ctypes.Structure):
    _fields_ = [("nk", ctypes.c_int64), ("mesh", ctypes.c_int64 * 3), ("nao", ctypes.c_int64),
                ("nso", ctypes.c_int64), ("ns", ctypes.c_int64), ("NQ", ctypes.c_int64),
                ("nts", ctypes.c_int64), ("nw_b", ctypes.c_int64), ("beta", ctypes.c_double)]


class DysonDims(ctypes.Structure):
    _fields_ = [("nts", ctypes.c_int64), ("nw", ctypes.c_int64), ("ns", ctypes.c_int64), ("nk", ctypes.c_int64),
                ("ink", ctypes.c_int64), ("nao", ctypes.c_int64), ("nso", ctypes.c_int64)]


_lib = None


def load() -> ctypes.CDLL:
    """Load libgwhip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python green-mbpt_amd/build.py` "
                           "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    lib.gwhip_create.argtypes = [ctypes.POINTER(Dims), ctypes.c_int, ctypes.POINTER(P)]
    lib.gwhip_create.restype = ctypes.c_int
    lib.gwhip_destroy.argtypes = [P]
    lib.gwhip_destroy.restype = None
    lib.gwhip_last_error.argtypes = [P]
    lib.gwhip_last_error.restype = ctypes.c_char_p
    lib.gwhip_set_grid.argtypes = [P, P, P]
    lib.gwhip_upload_V.argtypes = [P, P]
    lib.gwhip_set_kq_maps.argtypes = [P, P, P]
    lib.gwhip_set_split_slices.argtypes = [P, ctypes.c_int]
    lib.gwhip_generate_V.argtypes = [P, ctypes.c_uint64, ctypes.c_double]
    lib.gwhip_eval_sigma.argtypes = [P, P, P, ctypes.c_int]
    lib.gwhip_eval_sigma_device.argtypes = [P, P, P, ctypes.c_int]
    lib.gwhip_eval_sigma_device_noalloc.argtypes = [P, P, P, ctypes.c_int]
    lib.gwhip_get_P_tilde.argtypes = [P, P]
    lib.gwhip_timings.argtypes = [P, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    lib.gwhip_kernel_stats.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int64),
                                       ctypes.POINTER(ctypes.c_double)]
    lib.gwhip_device_bytes.argtypes = [P]
    lib.gwhip_device_bytes.restype = ctypes.c_size_t
    lib.gwhip_debug_zgemm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, P, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int]
    lib.gwhip_debug_stage.argtypes = [P, P, ctypes.c_int]
    lib.gwhip_debug_eval_q.argtypes = [P, P, ctypes.c_int64, ctypes.c_int, P, P, P]
    lib.gwhip_set_leakage_capture.argtypes = [P, ctypes.c_int]
    lib.gwhip_get_leakage_inputs.argtypes = [P, P, P]
    lib.gwhip_debug_zgemm_split.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, P, ctypes.c_int, ctypes.c_int]
    lib.gwhip_autotune.argtypes = [P, P, ctypes.c_int]
    IP = ctypes.POINTER(ctypes.c_int)
    lib.gwhip_get_tuning.argtypes = [P, IP, ctypes.c_int]
    lib.gwhip_set_tuning.argtypes = [P, IP, ctypes.c_int]
    lib.gwhip_batch_create.argtypes = [ctypes.c_int, ctypes.POINTER(Dims), ctypes.c_int, ctypes.POINTER(P)]
    lib.gwhip_batch_destroy.argtypes = [P]
    lib.gwhip_batch_destroy.restype = None
    lib.gwhip_batch_last_error.argtypes = [P]
    lib.gwhip_batch_last_error.restype = ctypes.c_char_p
    lib.gwhip_batch_system.argtypes = [P, ctypes.c_int]
    lib.gwhip_batch_system.restype = P
    lib.gwhip_batch_autotune.argtypes = [P, ctypes.POINTER(P), ctypes.c_int]
    lib.gwhip_batch_eval_device.argtypes = [P, ctypes.POINTER(P), ctypes.POINTER(P), ctypes.c_int]
    lib.gwhip_batch_stats.argtypes = [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64),
                                      ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_size_t)]
    lib.gwhip_set_vjp_store.argtypes = [P, ctypes.c_int]
    lib.gwhip_vjp.argtypes = [P, P, P, P, ctypes.c_int]
    lib.gwhip_vjp_device.argtypes = [P, P, P, P, ctypes.c_int]
    lib.gwhip_vjp_ms.argtypes = [P]
    lib.gwhip_vjp_ms.restype = ctypes.c_double
    lib.gwhip_hf.argtypes = [P, P, P, ctypes.c_double, P]
    lib.gwhip_hf_device.argtypes = [P, P, P, ctypes.c_double, P]
    lib.gwhip_hf_ms.argtypes = [P]
    lib.gwhip_hf_ms.restype = ctypes.c_double
    lib.gwhip_hf_exchange_stats.argtypes = [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]
    lib.gwhip_set_k_symmetry.argtypes = [P, ctypes.c_int64, P, P, P, P]
    lib.gwhip_set_q_symmetry.argtypes = [P, ctypes.c_int64, P, P, P, P]
    lib.gwhip_eval_sigma_ibz.argtypes = [P, P, P, ctypes.c_int]
    lib.gwhip_eval_sigma_ibz_device.argtypes = [P, P, P, ctypes.c_int]
    lib.gwhip_hf_ibz.argtypes = [P, P, P, ctypes.c_double, P]
    lib.gwhip_hf_ibz_device.argtypes = [P, P, P, ctypes.c_double, P]
    lib.gwhip_get_dims.argtypes = [P, ctypes.POINTER(Dims)]
    lib.gwhip_set_q0_extrapolation.argtypes = [P, ctypes.c_int64, P, P, P, ctypes.c_double, ctypes.c_int64, P]
    lib.gwhip_get_q0_eps.argtypes = [P, P]
    lib.gwhip_dyson_create.argtypes = [ctypes.POINTER(DysonDims), ctypes.c_int, ctypes.POINTER(P)]
    lib.gwhip_dyson_destroy.argtypes = [P]
    lib.gwhip_dyson_destroy.restype = None
    lib.gwhip_dyson_last_error.argtypes = [P]
    lib.gwhip_dyson_last_error.restype = ctypes.c_char_p
    lib.gwhip_dyson_set_grid.argtypes = [P, P, P, P]
    lib.gwhip_dyson_set_system.argtypes = [P, P, P, P]
    lib.gwhip_dyson_compute_G.argtypes = [P, P, P, ctypes.c_double, P]
    lib.gwhip_dyson_compute_G_device.argtypes = [P, P, P, ctypes.c_double, P]
    DP = ctypes.POINTER(ctypes.c_double)
    lib.gwhip_dyson_find_mu.argtypes = [P, P, P, ctypes.c_double, ctypes.c_double, ctypes.c_double, DP, DP]
    lib.gwhip_dyson_find_mu_device.argtypes = [P, P, P, ctypes.c_double, ctypes.c_double, ctypes.c_double, DP, DP]
    lib.gwhip_dyson_number_of_electrons.argtypes = [P, P, P, ctypes.c_double, DP]
    lib.gwhip_dyson_energy.argtypes = [P, P, P, P, DP]
    lib.gwhip_dyson_energy_device.argtypes = [P, P, P, P, DP]
    for f in ("gwhip_debug_zgemm_split", "gwhip_set_split_slices", "gwhip_set_kq_maps", "gwhip_debug_zgemm", "gwhip_debug_eval_q",
              "gwhip_set_leakage_capture", "gwhip_get_leakage_inputs", "gwhip_debug_stage", "gwhip_set_grid", "gwhip_upload_V",
              "gwhip_generate_V", "gwhip_eval_sigma", "gwhip_eval_sigma_device", "gwhip_get_P_tilde", "gwhip_timings",
              "gwhip_kernel_stats", "gwhip_set_vjp_store", "gwhip_vjp", "gwhip_vjp_device", "gwhip_autotune",
              "gwhip_get_tuning", "gwhip_set_tuning", "gwhip_batch_create", "gwhip_batch_autotune",
              "gwhip_batch_eval_device", "gwhip_batch_stats", "gwhip_set_k_symmetry", "gwhip_set_q_symmetry",
              "gwhip_eval_sigma_ibz",
              "gwhip_eval_sigma_ibz_device", "gwhip_hf_ibz", "gwhip_hf_ibz_device", "gwhip_get_dims",
              "gwhip_set_q0_extrapolation", "gwhip_get_q0_eps", "gwhip_dyson_create", "gwhip_dyson_set_grid",
              "gwhip_dyson_set_system", "gwhip_dyson_compute_G", "gwhip_dyson_compute_G_device", "gwhip_dyson_find_mu",
              "gwhip_dyson_find_mu_device", "gwhip_dyson_number_of_electrons", "gwhip_dyson_energy",
              "gwhip_dyson_energy_device"):
        getattr(lib, f).restype = ctypes.c_int
    _lib = lib
    return lib


def header_functions(header: str = HEADER) -> list[str]:
    """Names of every function a C header declares."""
    text = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(gwhip_\w+)\s*\(", text, re.M)))


_h5 = None


def load_h5() -> ctypes.CDLL:
    """Load libgwhip_h5.so, the HDF5 integral reader (raises if it has not been built).  The image's
    libhdf5 is loaded first by its full path, so the reader finds it whatever the run path says."""
    global _h5
    if _h5 is not None:
        return _h5
    if not os.path.exists(H5_LIB_PATH):
        raise RuntimeError(f"{H5_LIB_PATH} is missing: build it with `python green-mbpt_amd/build.py`")
    load()
    libdir = os.path.join(HDF5_ROOT, "lib")
    for name in sorted(os.listdir(libdir)) if os.path.isdir(libdir) else []:
        if name.startswith("libhdf5.so.") and name.count(".") == 2:
            ctypes.CDLL(os.path.join(libdir, name), mode=ctypes.RTLD_GLOBAL)
            break
    h5 = ctypes.CDLL(H5_LIB_PATH)
    P, I64P = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)
    h5.gwhip_h5_check_version.argtypes = [ctypes.c_char_p]
    h5.gwhip_h5_read_meta.argtypes = [ctypes.c_char_p, I64P]
    h5.gwhip_h5_read_packed.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int64, ctypes.c_int64,
                                        ctypes.c_int64, P]
    h5.gwhip_h5_read_k_symmetry.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int64, I64P, P, P, P, P, P]
    h5.gwhip_h5_read_q_symmetry.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int64, I64P, P, P, P, P]
    h5.gwhip_h5_kq_maps.argtypes = [P, ctypes.c_int64, P, P]
    h5.gwhip_h5_load.argtypes = [P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    h5.gwhip_h5_load_q0.argtypes = [P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int64, P,
                                    ctypes.POINTER(ctypes.c_int)]
    h5.gwhip_h5_last_error.restype = ctypes.c_char_p
    _h5 = h5
    return h5


def _h5check(st: int):
    if st != OK:
        _raise(st, load_h5().gwhip_h5_last_error().decode())


def h5_check_version(version: str) -> int:
    """1 if version >= 1.0.0, 0 if older, -1 if malformed (CheckVersion)."""
    return int(load_h5().gwhip_h5_check_version(version.encode()))


def h5_read_meta(df_path: str) -> int:
    """chunk_size of an integral directory after the version check."""
    cs = ctypes.c_int64()
    _h5check(load_h5().gwhip_h5_read_meta(df_path.encode(), ctypes.byref(cs)))
    return cs.value


def h5_read_packed(df_path: str, input_h5: str, nk: int, nao: int, NQ: int) -> np.ndarray:
    """Every k pair k1 >= k2, packed [nk(nk+1)/2, NQ, nao, nao] complex128."""
    out = np.empty((nk * (nk + 1) // 2, NQ, nao, nao), np.complex128)
    _h5check(load_h5().gwhip_h5_read_packed(df_path.encode(), input_h5.encode(), nk, nao, NQ, _ptr(out)))
    return out


def h5_read_k_symmetry(input_h5: str, nk: int, nso: int) -> dict:
    """/symmetry/k tables: ink, bz2ibz, ibz2bz, U, tr_conj, mesh_scaled."""
    ink = ctypes.c_int64()
    b = np.empty(nk, np.int64)
    i = np.empty(nk, np.int64)
    t = np.empty(nk, np.int64)
    U = np.empty((nk, nso, nso), np.complex128)
    m = np.empty((nk, 3), np.float64)
    _h5check(load_h5().gwhip_h5_read_k_symmetry(input_h5.encode(), nk, nso, ctypes.byref(ink), _ptr(b), _ptr(i),
                                                _ptr(U), _ptr(t), _ptr(m)))
    return dict(ink=ink.value, bz2ibz=b, ibz2bz=i[:ink.value].copy(), U=U, tr_conj=t, mesh_scaled=m)


def h5_read_q_symmetry(input_h5: str, nq: int, NQ: int) -> dict:
    """/symmetry/q tables: inq, bz2ibz, ibz2bz, U (k_sym_transform_p0), tr_conj."""
    inq = ctypes.c_int64()
    b = np.empty(nq, np.int64)
    i = np.empty(nq, np.int64)
    t = np.empty(nq, np.int64)
    U = np.empty((nq, NQ, NQ), np.complex128)
    _h5check(load_h5().gwhip_h5_read_q_symmetry(input_h5.encode(), nq, NQ, ctypes.byref(inq), _ptr(b), _ptr(i),
                                                _ptr(U), _ptr(t)))
    return dict(inq=inq.value, bz2ibz=b, ibz2bz=i[:inq.value].copy(), U=U, tr_conj=t)


def h5_kq_maps(mesh_scaled: np.ndarray):
    """(kplus[q][k], kminus[q][k]) from fractional k coordinates."""
    x = np.ascontiguousarray(mesh_scaled, np.float64)
    nk = x.shape[0]
    kp = np.empty((nk, nk), np.int32)
    km = np.empty((nk, nk), np.int32)
    _h5check(load_h5().gwhip_h5_kq_maps(_ptr(x), nk, _ptr(kp), _ptr(km)))
    return kp, km


def debug_zgemm(A: np.ndarray, B: np.ndarray, conjA: bool = False, conjB: bool = False, shape: int = 0) -> np.ndarray:
    """The library's FP64 MFMA GEMM on host arrays (test hook)."""
    lib = load()
    a = np.ascontiguousarray(A, dtype=np.complex128)
    b = np.ascontiguousarray(B, dtype=np.complex128)
    M, K = a.shape
    K2, N = b.shape
    assert K == K2
    c = np.empty((M, N), dtype=np.complex128)
    st = lib.gwhip_debug_zgemm(M, N, K, a.ctypes.data, b.ctypes.data, c.ctypes.data, int(conjA), int(conjB), shape)
    if st != OK:
        raise GwhipError(st, "debug_zgemm failed")
    return c


def debug_zgemm_split(A: np.ndarray, B: np.ndarray, slices: int = 4, tile: int = 2) -> np.ndarray:
    """The split-precision (int8 slice) GEMM on host arrays with block tile `tile` (2 or 4; test hook)."""
    lib = load()
    a = np.ascontiguousarray(A, dtype=np.complex128)
    b = np.ascontiguousarray(B, dtype=np.complex128)
    M, K = a.shape
    N = b.shape[1]
    c = np.empty((M, N), dtype=np.complex128)
    st = lib.gwhip_debug_zgemm_split(M, N, K, a.ctypes.data, b.ctypes.data, c.ctypes.data, slices, tile)
    if st != OK:
        raise GwhipError(st, "debug_zgemm_split failed")
    return c


_dbg = None


def load_debug() -> ctypes.CDLL:
    """Load libgwhip_debug.so, the throughput probes (include/gwhip_debug.h)."""
    global _dbg
    if _dbg is not None:
        return _dbg
    if not os.path.exists(DEBUG_LIB_PATH):
        raise RuntimeError(f"{DEBUG_LIB_PATH} is missing: build it with `python green-mbpt_amd/build.py`")
    d = ctypes.CDLL(DEBUG_LIB_PATH)
    d.gwhip_debug_gemm_bench.argtypes = [ctypes.c_int] * 7 + [ctypes.POINTER(ctypes.c_double)]
    d.gwhip_debug_mfma_peak.argtypes = [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_double)]
    d.gwhip_debug_gemm_bench.restype = ctypes.c_int
    d.gwhip_debug_mfma_peak.restype = ctypes.c_int
    _dbg = d
    return d


def debug_gemm_bench(which: int, NQ: int, nk: int, nao: int, batch: int, shape: int = 0, reps: int = 3) -> float:
    """Mean ms per launch of the FP64 GEMM in the Sigma (0) or P0 (1) layout (probe library)."""
    ms = ctypes.c_double()
    st = load_debug().gwhip_debug_gemm_bench(which, NQ, nk, nao, batch, shape, reps, ctypes.byref(ms))
    if st != OK:
        raise GwhipError(st, "debug_gemm_bench failed")
    return ms.value


def debug_mfma4_layout() -> np.ndarray:
    """[3][64] outputs of the v_mfma_f64_4x4x4_4b layout probe (test hook)."""
    d = load_debug()
    d.gwhip_debug_mfma4_layout.argtypes = [ctypes.POINTER(ctypes.c_double)]
    d.gwhip_debug_mfma4_layout.restype = ctypes.c_int
    out = np.zeros(3 * 64)
    st = d.gwhip_debug_mfma4_layout(out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    if st != OK:
        raise GwhipError(st, "mfma4 layout probe failed")
    return out.reshape(3, 64)


def debug_mfma_peak(chains: int, waves_per_simd: int = 2, iters: int = 4096) -> float:
    """Register-only throughput probe (probe library): TFLOP/s or TOPS."""
    v = ctypes.c_double()
    st = load_debug().gwhip_debug_mfma_peak(chains, waves_per_simd, iters, ctypes.byref(v))
    if st != OK:
        raise GwhipError(st, "debug_mfma_peak failed")
    return v.value


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def default_mesh(nk, mesh=None):
    if mesh is not None:
        return tuple(mesh)
    n = round(nk ** (1.0 / 3.0))
    return (n, n, n) if n ** 3 == nk else (nk, 1, 1)


class GwHipKernel:
    """One device-resident GW kernel instance (dims fixed at construction)."""

    def __init__(self, nk, nao, nso, ns, NQ, nts, nw_b, beta, mesh=None, device=0, _borrowed=None):
        self.lib = load()
        mesh = default_mesh(nk, mesh)
        self.dims = dict(nk=nk, nao=nao, nso=nso, ns=ns, NQ=NQ, nts=nts, nw_b=nw_b, beta=beta, mesh=tuple(mesh))
        if _borrowed is not None:  # a system of a Batch: the batch owns the handle
            self.h, self._owner = _borrowed, False
            return
        self._owner = True
        d = Dims(nk, (ctypes.c_int64 * 3)(*mesh), nao, nso, ns, NQ, nts, nw_b, beta)
        h = ctypes.c_void_p()
        st = self.lib.gwhip_create(ctypes.byref(d), device, ctypes.byref(h))
        if st != OK:
            msg = self.lib.gwhip_last_error(None).decode()
            raise (WrongGrid if st == ERR_WRONG_GRID else GwhipError)(st, msg)
        self.h = h

    def _check(self, st: int):
        if st != OK:
            _raise(st, self.lib.gwhip_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "_owner", True):
                self.lib.gwhip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_grid(self, T_bt: np.ndarray, T_tb: np.ndarray):
        a = np.ascontiguousarray(T_bt, dtype=np.complex128)
        b = np.ascontiguousarray(T_tb, dtype=np.complex128)
        assert a.shape == (self.dims["nw_b"], self.dims["nts"]) and b.shape == (self.dims["nts"], self.dims["nw_b"])
        self._check(self.lib.gwhip_set_grid(self.h, _ptr(a), _ptr(b)))

    def set_kq_maps(self, kplus: np.ndarray, kminus: np.ndarray):
        """Explicit k + q / k - q tables [nq][nk] (instead of the Monkhorst-Pack mesh maps)."""
        a = np.ascontiguousarray(kplus, dtype=np.int32)
        b = np.ascontiguousarray(kminus, dtype=np.int32)
        nk = self.dims["nk"]
        assert a.shape == (nk, nk) and b.shape == (nk, nk)
        self._check(self.lib.gwhip_set_kq_maps(self.h, _ptr(a), _ptr(b)))

    def set_split_slices(self, ns: int):
        self._check(self.lib.gwhip_set_split_slices(self.h, ns))

    def upload_V(self, Vp: np.ndarray):
        nk, NQ, nao = self.dims["nk"], self.dims["NQ"], self.dims["nao"]
        v = np.ascontiguousarray(Vp, dtype=np.complex128)
        assert v.shape == (nk * (nk + 1) // 2, NQ, nao, nao), v.shape
        self._check(self.lib.gwhip_upload_V(self.h, _ptr(v)))

    def generate_V(self, seed: int, sigma: float = 0.1):
        self._check(self.lib.gwhip_generate_V(self.h, ctypes.c_uint64(seed), sigma))

    def sigma_shape(self):
        d = self.dims
        return (d["nts"], d["ns"], d["nk"], d["nso"], d["nso"])

    def solve(self, G: np.ndarray, precision: int = PREC_FP64) -> np.ndarray:
        g = np.ascontiguousarray(G, dtype=np.complex128)
        assert g.shape == self.sigma_shape(), (g.shape, self.sigma_shape())
        S = np.empty_like(g)
        self._check(self.lib.gwhip_eval_sigma(self.h, _ptr(g), _ptr(S), precision))
        return S

    def solve_device(self, G_ptr: int, S_ptr: int, precision: int = PREC_FP64):
        self._check(self.lib.gwhip_eval_sigma_device(self.h, ctypes.c_void_p(G_ptr), ctypes.c_void_p(S_ptr),
                                                     precision))

    def solve_device_noalloc(self, G_ptr: int, S_ptr: int, precision: int = PREC_FP64):
        """gwhip_eval_sigma_device_noalloc: the same evaluation inside the library's fixed device
        workspace (no allocation)."""
        self._check(self.lib.gwhip_eval_sigma_device_noalloc(self.h, ctypes.c_void_p(G_ptr), ctypes.c_void_p(S_ptr),
                                                             precision))

    def set_vjp_store(self, keep: bool):
        self._check(self.lib.gwhip_set_vjp_store(self.h, 1 if keep else 0))

    def vjp(self, G: np.ndarray, Sbar: np.ndarray, mode: int = VJP_RECOMPUTE) -> np.ndarray:
        """G_bar with Re<Sbar, dSigma> = Re<G_bar, dG> (host arrays)."""
        g = np.ascontiguousarray(G, dtype=np.complex128)
        sb = np.ascontiguousarray(Sbar, dtype=np.complex128)
        assert g.shape == self.sigma_shape() and sb.shape == g.shape
        out = np.empty_like(g)
        self._check(self.lib.gwhip_vjp(self.h, _ptr(g), _ptr(sb), _ptr(out), mode))
        return out

    def vjp_device(self, G_ptr: int, Sbar_ptr: int, Gbar_ptr: int, mode: int = VJP_RECOMPUTE):
        self._check(self.lib.gwhip_vjp_device(self.h, ctypes.c_void_p(G_ptr), ctypes.c_void_p(Sbar_ptr),
                                              ctypes.c_void_p(Gbar_ptr), mode))

    def autotune(self, G_ptr: int, precision: int = PREC_FP64):
        self._check(self.lib.gwhip_autotune(self.h, ctypes.c_void_p(G_ptr), precision))

    def tuning(self) -> dict:
        v = (ctypes.c_int * len(TUNING_KEYS))()
        self._check(self.lib.gwhip_get_tuning(self.h, v, len(TUNING_KEYS)))
        return dict(zip(TUNING_KEYS, list(v)))

    def set_tuning(self, **kw):
        """Change any of TUNING_KEYS (the others keep their current values)."""
        t = self.tuning()
        unknown = set(kw) - set(TUNING_KEYS)
        if unknown:
            raise KeyError(sorted(unknown))
        t.update(kw)
        v = (ctypes.c_int * len(TUNING_KEYS))(*[t[k] for k in TUNING_KEYS])
        self._check(self.lib.gwhip_set_tuning(self.h, v, len(TUNING_KEYS)))

    def vjp_ms(self) -> float:
        return float(self.lib.gwhip_vjp_ms(self.h))

    def set_k_symmetry(self, ink: int, bz2ibz, ibz2bz, U, tr_conj):
        """Irreducible-wedge tables (input.h5:/symmetry/k); U: [nk, nso, nso] complex."""
        d = self.dims
        b = np.ascontiguousarray(bz2ibz, np.int64)
        i = np.ascontiguousarray(ibz2bz, np.int64)
        u = np.ascontiguousarray(U, np.complex128)
        t = np.ascontiguousarray(tr_conj, np.int64)
        assert b.shape == (d["nk"],) and i.shape == (ink,) and t.shape == (d["nk"],)
        assert u.shape == (d["nk"], d["nso"], d["nso"])
        self._check(self.lib.gwhip_set_k_symmetry(self.h, ink, _ptr(b), _ptr(i), _ptr(u), _ptr(t)))
        self.dims["ink"] = ink

    def set_q_symmetry(self, inq: int, bz2ibz=None, ibz2bz=None, U=None, tr_conj=None):
        """q-symmetry tables (input.h5:/symmetry/q); U: [nq, NQ, NQ] complex.  With the k symmetry
        set, solve_ibz then runs the star-of-q evaluation; inq = 0 switches it off."""
        d = self.dims
        if inq == 0:
            self._check(self.lib.gwhip_set_q_symmetry(self.h, 0, None, None, None, None))
            return
        b = np.ascontiguousarray(bz2ibz, np.int64)
        i = np.ascontiguousarray(ibz2bz, np.int64)
        u = np.ascontiguousarray(U, np.complex128)
        t = np.ascontiguousarray(tr_conj, np.int64)
        assert b.shape == (d["nk"],) and i.shape == (inq,) and t.shape == (d["nk"],)
        assert u.shape == (d["nk"], d["NQ"], d["NQ"])
        self._check(self.lib.gwhip_set_q_symmetry(self.h, inq, _ptr(b), _ptr(i), _ptr(u), _ptr(t)))

    def load_files(self, df_path: str, input_h5: str, with_symmetry: bool = True):
        """k +- q maps, integrals and (optionally) the k and q symmetry from green-mbpt's files."""
        st = load_h5().gwhip_h5_load(self.h, df_path.encode(), input_h5.encode(), 1 if with_symmetry else 0)
        _h5check(st)
        if with_symmetry:
            self.dims["ink"] = h5_read_k_symmetry(input_h5, self.dims["nk"], self.dims["nso"])["ink"]

    def set_q0_extrapolation(self, q_ir, AqQ, q_abs, madelung: float, S):
        """q -> 0 (EXTRAPOLATE) correction; S [ns, nk or ink, nso, nso]; an empty q_ir switches it off."""
        qi = np.ascontiguousarray(q_ir, np.int64)
        A = np.ascontiguousarray(AqQ, np.complex128)
        qa = np.ascontiguousarray(q_abs, np.float64)
        Sm = np.ascontiguousarray(S, np.complex128)
        self._check(self.lib.gwhip_set_q0_extrapolation(self.h, qi.shape[0], _ptr(qi), _ptr(A), _ptr(qa),
                                                        float(madelung), Sm.shape[1], _ptr(Sm)))

    def load_q0(self, df_path: str, input_h5: str, S) -> bool:
        """AqQ.h5 of an integral directory; False (correction off) when the file is absent."""
        Sm = np.ascontiguousarray(S, np.complex128)
        en = ctypes.c_int()
        _h5check(load_h5().gwhip_h5_load_q0(self.h, df_path.encode(), input_h5.encode(), Sm.shape[1], _ptr(Sm),
                                            ctypes.byref(en)))
        return bool(en.value)

    def q0_eps(self) -> np.ndarray:
        """eps_q0 on the tau grid of the last corrected evaluation [nts]."""
        out = np.empty(self.dims["nts"], np.complex128)
        self._check(self.lib.gwhip_get_q0_eps(self.h, _ptr(out)))
        return out

    def ibz_shape(self, with_tau: bool = True):
        d = self.dims
        s = (d["ns"], d["ink"], d["nso"], d["nso"])
        return ((d["nts"],) + s) if with_tau else s

    def solve_ibz(self, G_ibz: np.ndarray, precision: int = PREC_FP64) -> np.ndarray:
        """Sigma on the irreducible k points from G on them (host arrays)."""
        g = np.ascontiguousarray(G_ibz, dtype=np.complex128)
        assert g.shape == self.ibz_shape(), (g.shape, self.ibz_shape())
        S = np.empty_like(g)
        self._check(self.lib.gwhip_eval_sigma_ibz(self.h, _ptr(g), _ptr(S), precision))
        return S

    def solve_ibz_device(self, G_ptr: int, S_ptr: int, precision: int = PREC_FP64):
        self._check(self.lib.gwhip_eval_sigma_ibz_device(self.h, ctypes.c_void_p(G_ptr), ctypes.c_void_p(S_ptr),
                                                         precision))

    def hf_ibz(self, dm: np.ndarray, S: np.ndarray, madelung: float) -> np.ndarray:
        """Fock contribution on the irreducible k points (host arrays [ns, ink, nso, nso])."""
        d = np.ascontiguousarray(dm, dtype=np.complex128)
        s = np.ascontiguousarray(S, dtype=np.complex128)
        assert d.shape == self.ibz_shape(False) and s.shape == d.shape, (d.shape, s.shape)
        F = np.empty_like(d)
        self._check(self.lib.gwhip_hf_ibz(self.h, _ptr(d), _ptr(s), float(madelung), _ptr(F)))
        return F

    def hf_shape(self):
        d = self.dims
        return (d["ns"], d["nk"], d["nso"], d["nso"])

    def hf(self, dm: np.ndarray, S: np.ndarray, madelung: float) -> np.ndarray:
        """Hartree-Fock Fock contribution [ns][nk][nso][nso] (host arrays), the counterpart of
        hf_scalar_cpu_kernel / hf_x2c_cpu_kernel::solve (src/hf_cpu_kernels.cpp:24-279)."""
        d = np.ascontiguousarray(dm, dtype=np.complex128)
        s = np.ascontiguousarray(S, dtype=np.complex128)
        assert d.shape == self.hf_shape() and s.shape == d.shape, (d.shape, s.shape, self.hf_shape())
        F = np.empty_like(d)
        self._check(self.lib.gwhip_hf(self.h, _ptr(d), _ptr(s), float(madelung), _ptr(F)))
        return F

    def hf_device(self, dm_ptr: int, S_ptr: int, madelung: float, F_ptr: int):
        self._check(self.lib.gwhip_hf_device(self.h, ctypes.c_void_p(dm_ptr), ctypes.c_void_p(S_ptr),
                                             float(madelung), ctypes.c_void_p(F_ptr)))

    def hf_ms(self) -> float:
        return float(self.lib.gwhip_hf_ms(self.h))

    def hf_exchange_stats(self):
        """(milliseconds, launches) of the exchange-term kernel in the last hf call."""
        ms, n = ctypes.c_double(), ctypes.c_int64()
        self._check(self.lib.gwhip_hf_exchange_stats(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def run_stages(self, G: np.ndarray, last_stage: int):
        g = np.ascontiguousarray(G, dtype=np.complex128)
        self._check(self.lib.gwhip_debug_stage(self.h, _ptr(g), last_stage))

    def P_tilde(self) -> np.ndarray:
        d = self.dims
        P = np.empty((d["nk"], d["nts"], d["NQ"], d["NQ"]), dtype=np.complex128)
        self._check(self.lib.gwhip_get_P_tilde(self.h, _ptr(P)))
        return P

    def eval_q(self, G: np.ndarray, q: int, precision: int = PREC_FP64, want=("P0", "P_tilde", "Sigma_q")) -> dict:
        """P0(tau, q), P-tilde(tau, q) and q's Sigma contribution (1/nk included) from one q alone."""
        d = self.dims
        g = np.ascontiguousarray(G, dtype=np.complex128)
        pq = (d["nts"], d["NQ"], d["NQ"])
        out = {"P0": np.empty(pq, complex) if "P0" in want else None,
               "P_tilde": np.empty(pq, complex) if "P_tilde" in want else None,
               "Sigma_q": np.empty(self.sigma_shape(), complex) if "Sigma_q" in want else None}
        ptr = {n: (_ptr(a) if a is not None else None) for n, a in out.items()}
        self._check(self.lib.gwhip_debug_eval_q(self.h, _ptr(g), int(q), int(precision), ptr["P0"], ptr["P_tilde"],
                                                ptr["Sigma_q"]))
        return {n: a for n, a in out.items() if a is not None}

    def set_leakage_capture(self, on: bool):
        self._check(self.lib.gwhip_set_leakage_capture(self.h, 1 if on else 0))

    def leakage_inputs(self):
        """(P0, P-tilde) of the first q of the last evaluation, [nts, NQ, NQ] each."""
        d = self.dims
        P0 = np.empty((d["nts"], d["NQ"], d["NQ"]), complex)
        Pt = np.empty_like(P0)
        self._check(self.lib.gwhip_get_leakage_inputs(self.h, _ptr(P0), _ptr(Pt)))
        return P0, Pt

    def timings(self) -> dict:
        buf = (ctypes.c_double * len(REGIONS))()
        n = self.lib.gwhip_timings(self.h, buf, len(REGIONS))
        return {REGIONS[i]: buf[i] for i in range(n)}

    def kernel_stats(self) -> dict:
        out = {}
        for i, name in enumerate(KSTATS):
            n = ctypes.c_int64()
            ms = ctypes.c_double()
            self._check(self.lib.gwhip_kernel_stats(self.h, i, ctypes.byref(n), ctypes.byref(ms)))
            out[name] = (n.value, ms.value)
        return out

    def device_bytes(self) -> int:
        return int(self.lib.gwhip_device_bytes(self.h))


class Batch:
    """Several systems evaluated by one captured hipGraph (gwhip_batch_*)."""

    def __init__(self, systems: list[dict], device: int = 0):
        """systems: dicts with nk, nao, nso, ns, NQ, nts, nw_b, beta and optional mesh."""
        self.lib = load()
        n = len(systems)
        arr = (Dims * n)()
        for i, s in enumerate(systems):
            mesh = default_mesh(s["nk"], s.get("mesh"))
            arr[i] = Dims(s["nk"], (ctypes.c_int64 * 3)(*mesh), s["nao"], s["nso"], s["ns"], s["NQ"], s["nts"],
                          s["nw_b"], s["beta"])
        b = ctypes.c_void_p()
        st = self.lib.gwhip_batch_create(n, arr, device, ctypes.byref(b))
        if st != OK:
            raise GwhipError(st, self.lib.gwhip_batch_last_error(None).decode())
        self.b = b
        self.systems = []
        for i, s in enumerate(systems):
            h = self.lib.gwhip_batch_system(self.b, i)
            self.systems.append(GwHipKernel(s["nk"], s["nao"], s["nso"], s["ns"], s["NQ"], s["nts"], s["nw_b"],
                                            s["beta"], mesh=s.get("mesh"), _borrowed=ctypes.c_void_p(h)))

    def _check(self, st: int):
        if st != OK:
            raise GwhipError(st, self.lib.gwhip_batch_last_error(self.b).decode())

    @staticmethod
    def _ptrs(ptrs):
        return (ctypes.c_void_p * len(ptrs))(*[ctypes.c_void_p(p) for p in ptrs])

    def autotune(self, G_ptrs, precision: int = PREC_FP64):
        self._check(self.lib.gwhip_batch_autotune(self.b, self._ptrs(G_ptrs), precision))

    def eval_device(self, G_ptrs, S_ptrs, precision: int = PREC_FP64):
        assert len(G_ptrs) == len(self.systems) == len(S_ptrs)
        self._check(self.lib.gwhip_batch_eval_device(self.b, self._ptrs(G_ptrs), self._ptrs(S_ptrs), precision))

    def stats(self) -> dict:
        ms, nodes, caps, ws = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_size_t()
        self._check(self.lib.gwhip_batch_stats(self.b, ctypes.byref(ms), ctypes.byref(nodes), ctypes.byref(caps),
                                               ctypes.byref(ws)))
        return dict(last_ms=ms.value, graph_nodes=nodes.value, captures=caps.value, workspace_bytes=ws.value)

    def close(self):
        if getattr(self, "b", None):
            for s in self.systems:
                s.h = None
            self.lib.gwhip_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Dyson:
    """Device Dyson step and chemical-potential search (``dyson<G, S1, St>``, ``src/dyson.cpp:19-236``).

    Arrays on the irreducible k points: H, S, Sigma1 [ns, ink, nso, nso]; Sigma(tau), G(tau)
    [nts, ns, ink, nso, nso].  As in the reference, Sigma1 and Sigma(tau) are made Hermitian in
    place (the host arrays passed in are overwritten with their Hermitian parts).
    """

    def __init__(self, nts, nw, ns, nk, ink, nao, nso, device=0):
        self.lib = load()
        self.dims = dict(nts=nts, nw=nw, ns=ns, nk=nk, ink=ink, nao=nao, nso=nso)
        d = DysonDims(nts, nw, ns, nk, ink, nao, nso)
        h = ctypes.c_void_p()
        st = self.lib.gwhip_dyson_create(ctypes.byref(d), device, ctypes.byref(h))
        if st != OK:
            _raise(st, self.lib.gwhip_dyson_last_error(None).decode())
        self.h = h

    def _check(self, st):
        if st != OK:
            _raise(st, self.lib.gwhip_dyson_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.gwhip_dyson_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_grid(self, T_ft, T_tf, wn):
        self._T = (np.ascontiguousarray(T_ft, np.complex128), np.ascontiguousarray(T_tf, np.complex128),
                   np.ascontiguousarray(wn, np.float64))
        self._check(self.lib.gwhip_dyson_set_grid(self.h, *[_ptr(a) for a in self._T]))

    def set_system(self, H, S, weights=None):
        h = np.ascontiguousarray(H, np.complex128)
        s = np.ascontiguousarray(S, np.complex128)
        w = None if weights is None else np.ascontiguousarray(weights, np.float64)
        self._check(self.lib.gwhip_dyson_set_system(self.h, _ptr(h), _ptr(s), None if w is None else _ptr(w)))

    def _sig(self, Sigma1, Sigma_tau):
        s1 = np.ascontiguousarray(Sigma1, np.complex128).copy()
        st = np.ascontiguousarray(Sigma_tau, np.complex128).copy()
        return s1, st

    def compute_G(self, Sigma1, Sigma_tau, mu: float) -> np.ndarray:
        s1, st = self._sig(Sigma1, Sigma_tau)
        G = np.empty_like(st)
        self._check(self.lib.gwhip_dyson_compute_G(self.h, _ptr(s1), _ptr(st), float(mu), _ptr(G)))
        return G

    def compute_G_device(self, s1_ptr: int, st_ptr: int, mu: float, G_ptr: int):
        self._check(self.lib.gwhip_dyson_compute_G_device(self.h, ctypes.c_void_p(s1_ptr), ctypes.c_void_p(st_ptr),
                                                          float(mu), ctypes.c_void_p(G_ptr)))

    def find_mu(self, Sigma1, Sigma_tau, mu0: float, nel: float, tol: float):
        """(electron count found, mu)."""
        s1, st = self._sig(Sigma1, Sigma_tau)
        mu, nf = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.gwhip_dyson_find_mu(self.h, _ptr(s1), _ptr(st), float(mu0), float(nel), float(tol),
                                                 ctypes.byref(mu), ctypes.byref(nf)))
        return nf.value, mu.value

    def number_of_electrons(self, Sigma1, Sigma_tau, mu: float) -> float:
        s1, st = self._sig(Sigma1, Sigma_tau)
        n = ctypes.c_double()
        self._check(self.lib.gwhip_dyson_number_of_electrons(self.h, _ptr(s1), _ptr(st), float(mu), ctypes.byref(n)))
        return n.value

    def energy(self, G_tau, Sigma1, Sigma_tau):
        """compute_energy on the device: (E_1b, E_HF, E_corr) (``src/common_utils.cpp:5-55``)."""
        g = np.ascontiguousarray(G_tau, dtype=np.complex128)
        s1 = np.ascontiguousarray(Sigma1, dtype=np.complex128)
        st = np.ascontiguousarray(Sigma_tau, dtype=np.complex128)
        e = (ctypes.c_double * 3)()
        self._check(self.lib.gwhip_dyson_energy(self.h, _ptr(g), _ptr(s1), _ptr(st), e))
        return float(e[0]), float(e[1]), float(e[2])

    def energy_device(self, G_ptr: int, s1_ptr: int, st_ptr: int):
        e = (ctypes.c_double * 3)()
        self._check(self.lib.gwhip_dyson_energy_device(self.h, ctypes.c_void_p(G_ptr), ctypes.c_void_p(s1_ptr),
                                                       ctypes.c_void_p(st_ptr), e))
        return float(e[0]), float(e[1]), float(e[2])
