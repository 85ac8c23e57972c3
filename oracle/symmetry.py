"""CPU restatement of the k-point symmetry maps the GW and HF kernels use -- TEST INFRASTRUCTURE ONLY.

The reference calls green-symmetry (``brillouin_zone_utils``, a dependency that is not part of the
reference tree) for: the irreducible (IBZ) point of every k, ``value_AO`` (an irreducible-k object
carried to any k), ``full_to_ibz`` and the k +- q maps.  The input file holds everything these
need (``input.h5:/symmetry/k``): ``bz2ibz`` (the full-mesh index of each k's representative),
``ibz2bz`` (the representatives, in IBZ order), ``k_sym_transform_ao`` (one nao x nao AO-basis
matrix per k), ``tr_conj`` (1 where the symmetry includes time reversal) and ``mesh_scaled``
(fractional coordinates).  Uses: ``hf_cpu_kernels.cpp:107``, ``gw_cpu_kernel.cpp:177-178,336``.

Only ``tests/`` imports it.  The convention of ``value_AO`` here is the one the reference's nk = 27
HF and GW fixtures (``solvers_test.cpp:387,389``) reproduce.
"""
from __future__ import annotations

import numpy as np


def reduced_point(bz2ibz, ibz2bz, k: int) -> int:
    """IBZ ordinal of full-mesh point k."""
    rep = int(bz2ibz[k])
    pos = np.flatnonzero(np.asarray(ibz2bz) == rep)
    assert pos.size == 1
    return int(pos[0])


def value_AO(obj_ibz: np.ndarray, k: int, bz2ibz, ibz2bz, U, tr_conj) -> np.ndarray:
    """obj(k) from obj on the IBZ (``obj_ibz`` [ink, n, n]): U_k obj(ik) U_k^H, conjugated first
    when the symmetry includes time reversal; obj(ik) itself at a representative."""
    ik = reduced_point(bz2ibz, ibz2bz, k)
    x = obj_ibz[ik]
    if int(ibz2bz[ik]) == k:
        return x.copy()
    if tr_conj[k]:
        x = np.conj(x)
    u = U[k]
    return u @ x @ np.conj(u.T)


def unfold(obj_ibz: np.ndarray, bz2ibz, ibz2bz, U, tr_conj) -> np.ndarray:
    """``value_AO`` for every k: [..., ink, n, n] -> [..., nk, n, n]."""
    nk = len(bz2ibz)
    lead = obj_ibz.shape[:-3]
    flat = obj_ibz.reshape((-1,) + obj_ibz.shape[-3:])
    out = np.empty((flat.shape[0], nk) + obj_ibz.shape[-2:], np.complex128)
    for b in range(flat.shape[0]):
        for k in range(nk):
            out[b, k] = value_AO(flat[b], k, bz2ibz, ibz2bz, U, tr_conj)
    return out.reshape(lead + out.shape[1:])


def full_to_ibz(obj_full: np.ndarray, ibz2bz) -> np.ndarray:
    """[..., nk, n, n] -> [..., ink, n, n]: the representatives' entries."""
    return np.ascontiguousarray(np.take(obj_full, np.asarray(ibz2bz), axis=-3))


def kq_maps(mesh_scaled: np.ndarray, tol: float = 1e-6):
    """(kplus[q][k], kminus[q][k]): the mesh point at k + q and at k - q (mod reciprocal lattice),
    matched on fractional coordinates."""
    x = np.asarray(mesh_scaled, float)
    nk = x.shape[0]

    def find(y):
        d = x[None, :, :] - y[:, None, :]
        d -= np.round(d)
        hit = np.all(np.abs(d) < tol, axis=2)
        assert np.all(hit.sum(axis=1) == 1)
        return np.argmax(hit, axis=1)

    kp = np.empty((nk, nk), np.int32)
    km = np.empty((nk, nk), np.int32)
    for q in range(nk):
        kp[q] = find(x + x[q])
        km[q] = find(x - x[q])
    return kp, km
