"""CPU FP64 oracle for the Hartree-Fock kernel -- TEST INFRASTRUCTURE ONLY.

Restates ``hf_scalar_cpu_kernel::solve`` and ``hf_x2c_cpu_kernel::solve``
(``src/hf_cpu_kernels.cpp:24-146`` and ``:148-279``) in numpy for a full k mesh on one process
(``value_AO`` is the identity, ``full_point(ik) == ik``, the node split of Q and the allreduces
are identities).  Only ``tests/`` imports it; the device kernel never calls it.

Inputs: dm [ns, nk, nso, nso] (the caller's density matrix, ``hf_solver::solve`` forms it as
spin_prefactor * G(beta), ``src/hf_solver.cpp:9-16``), the packed integrals Vp
[nk(nk+1)/2, NQ, nao, nao] (``df_integral_t``), the overlap S [ns, nk, nso, nso] and the Madelung
constant.  Returns the Fock contribution [ns, nk, nso, nso] (X2C: [1, nk, nso, nso]).
"""
from __future__ import annotations

import numpy as np

from oracle import gw_oracle as go


def dm_from_G(G_beta: np.ndarray, ns: int, X2C: bool) -> np.ndarray:
    """``hf_solver``: dm = spin_prefactor * G(beta), -1 for ns = 2 or X2C, else -2."""
    return (-1.0 if (ns == 2 or X2C) else -2.0) * np.asarray(G_beta, np.complex128)


def hf_scalar(dm, Vp, S, madelung, nk, nao, ns, NQ):
    """``hf_scalar_cpu_kernel::solve`` (``src/hf_cpu_kernels.cpp:24-146``)."""
    F = np.zeros((ns, nk, nao, nao), np.complex128)
    # direct: (Q, 1) += (Q, ab) (ab, 1) with X1 = dm^T, summed over k' and spin (:45-59)
    upper = np.zeros(NQ, np.complex128)
    for ikp in range(nk):
        v = go.v_pair(Vp, ikp, ikp).reshape(NQ, nao * nao)
        for is_ in range(ns):
            upper += v @ dm[is_, ikp].T.reshape(nao * nao)
    upper /= nk
    for is_ in range(ns):
        for ik in range(nk):
            v = go.v_pair(Vp, ik, ik).reshape(NQ, nao * nao)
            F[is_, ik] += (upper @ v).reshape(nao, nao)          # (1, ij) = (1, Q) (Q, ij)  (:70-77)
    # exchange (:91-125): F(i, j) -= pref / nk * sum_{k', Q} (v dm)[Q, i, a] conj(v[Q, j, a])
    pref = 1.0 if ns == 2 else 0.5
    for is_ in range(ns):
        for ik in range(nk):
            for ikp in range(nk):
                v = go.v_pair(Vp, ik, ikp)                        # (Q, i, b)
                Y = v @ dm[is_, ikp]                              # (Q, i, a)
                F[is_, ik] -= pref * np.einsum("Qia,Qja->ij", Y, np.conj(v)) / nk
    # Ewald / Madelung (:128-136)
    for is_ in range(ns):
        for ik in range(nk):
            F[is_, ik] -= pref * madelung * S[is_, ik] @ dm[is_, ik] @ S[is_, ik]
    return F


def hf_x2c(dm, Vp, S, madelung, nk, nao, NQ):
    """``hf_x2c_cpu_kernel::solve`` (``src/hf_cpu_kernels.cpp:148-279``); dm, S: [1, nk, 2nao, 2nao]."""
    nso = 2 * nao
    F = np.zeros((1, nk, nso, nso), np.complex128)
    aa, bb = slice(0, nao), slice(nao, nso)
    upper = np.zeros(NQ, np.complex128)
    for ikp in range(nk):
        v = go.v_pair(Vp, ikp, ikp).reshape(NQ, nao * nao)
        X1 = (dm[0, ikp][aa, aa] + dm[0, ikp][bb, bb]).T
        upper += v @ X1.reshape(nao * nao)
    upper /= nk
    for ik in range(nk):
        v = go.v_pair(Vp, ik, ik).reshape(NQ, nao * nao)
        Fm = (upper @ v).reshape(nao, nao)
        F[0, ik][aa, aa] += Fm
        F[0, ik][bb, bb] += Fm
    for ik in range(nk):
        S_aa = S[0, ik][aa, aa]
        Faa = np.zeros((nao, nao), np.complex128)
        Fbb = np.zeros_like(Faa)
        Fab = np.zeros_like(Faa)
        for ikp in range(nk):
            v = go.v_pair(Vp, ik, ikp)
            d = dm[0, ikp]
            for blk, acc in (((aa, aa), Faa), ((bb, bb), Fbb), ((aa, bb), Fab)):
                Y = v @ d[blk]
                acc -= np.einsum("Qia,Qja->ij", Y, np.conj(v)) / nk
        F[0, ik][aa, aa] += Faa - madelung * S_aa @ dm[0, ik][aa, aa] @ S_aa
        F[0, ik][bb, bb] += Fbb - madelung * S_aa @ dm[0, ik][bb, bb] @ S_aa
        Fab -= madelung * S_aa @ dm[0, ik][aa, bb] @ S_aa
        F[0, ik][aa, bb] += Fab
        F[0, ik][bb, aa] = np.conj(F[0, ik][aa, bb]).T
    return F


def hf_fock(dm, Vp, S, madelung, nk, nao, nso, ns, NQ):
    """Dispatch as ``hf_kernel_factory`` does (X2C when nso != nao)."""
    if nso != nao:
        return hf_x2c(dm, Vp, S, madelung, nk, nao, NQ)
    return hf_scalar(dm, Vp, S, madelung, nk, nao, ns, NQ)
