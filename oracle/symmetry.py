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


# ---------------------------------------------------------------- star-of-q evaluation
def q_star(q_bz2ibz, q_ibz2bz, iq: int) -> list:
    """The full-mesh q points whose irreducible representative is q_ibz2bz[iq], ascending (the
    order of ``input.h5:/symmetry/q/stars/<iq>``)."""
    rep = int(q_ibz2bz[iq])
    return [int(q) for q in np.flatnonzero(np.asarray(q_bz2ibz) == rep)]


def p_bz_from_ibz(P: np.ndarray, U_q: np.ndarray, tr_conj: int) -> np.ndarray:
    """``eval_p0_bz_from_ibz`` (``src/gw_cpu_kernel.cpp:281-293``): P-tilde of the irreducible q
    carried to a q of its star, U_q P U_q^H in the auxiliary basis (U_q =
    ``/symmetry/q/k_sym_transform_p0[q]``), then conjugated when tr_conj[q] = 1."""
    Pq = U_q @ P @ np.conj(U_q.T)
    return np.conj(Pq) if tr_conj else Pq


def gw_sigma_star(G_ibz, Vp, T_bt, T_tb, sym, nk, nao, nso, ns, NQ, mesh, X2C=None, return_Pw=False):
    """``gw_cpu_kernel::solve`` on a symmetry-reduced mesh (``src/gw_cpu_kernel.cpp:31-134``):
    P0 and P-tilde for the irreducible q only (over all k1, with G carried to every k by
    ``value_AO``); Sigma for the irreducible k only, summed over every q of each star with the
    rotated P-tilde.  ``sym`` holds the k_* and q_* tables of ``input.h5:/symmetry``.

    G_ibz: [nts, ns, ink, nso, nso]; returns Sigma of the same shape (and the P-tilde(i Omega_b)
    stack [inq, nw_b, NQ, NQ] of the irreducible q when ``return_Pw``).
    """
    from oracle import gw_oracle as go

    G_ibz = np.asarray(G_ibz, dtype=np.complex128)
    nts = G_ibz.shape[0]
    if X2C is None:
        X2C = nao != nso
    if nts % 2 != 0:
        raise go.WrongGrid("Number of tau points should be even!!!")
    kb2i, ki2b = sym["k_bz2ibz"], sym["k_ibz2bz"]
    qb2i, qi2b = sym["q_bz2ibz"], sym["q_ibz2bz"]
    Uq, qtr = sym["q_k_sym_transform_p0"], sym["q_tr_conj"]
    ink, inq = len(ki2b), len(qi2b)
    G = unfold(G_ibz, kb2i, ki2b, sym["k_k_sym_transform_ao"], sym["k_tr_conj"])
    pns = nso // nao
    pseudo_ns = ns if not X2C else 4
    Sigma = np.zeros_like(G_ibz)
    Pws = []
    for iq in range(inq):
        q_ir = int(qi2b[iq])
        P0 = np.zeros((nts, NQ, NQ), dtype=np.complex128)
        go.eval_P0_tilde(P0, G, Vp, q_ir, nk, nao, nso, ns, NQ, X2C, mesh)
        go.symmetrize_P0(P0, nk)
        Pt, Pw = go.eval_P_tilde(P0, T_bt, T_tb, return_w=True)
        Pws.append(Pw)
        for k1 in range(ink):
            k1_ir = int(ki2b[k1])
            for q_deg in q_star(qb2i, qi2b, iq):
                k1mq = go.k_add(k1_ir, q_deg, mesh, -1)
                vm = go.v_pair(Vp, k1_ir, k1mq).reshape(NQ * nao, nao)
                for t in range(nts):
                    P = p_bz_from_ibz(Pt[t], Uq[q_deg], int(qtr[q_deg]))
                    for s in range(pseudo_ns):
                        is_ = s // (pns * pns)
                        a = (s // pns) % pns
                        b = s % pns
                        Gk = G[t, is_, k1mq][a * nao:(a + 1) * nao, b * nao:(b + 1) * nao]
                        Sm = go.selfenergy_contraction(Gk, vm, P, nao, NQ)
                        Sigma[t, is_, k1][a * nao:(a + 1) * nao, b * nao:(b + 1) * nao] -= Sm
    Sigma /= nk
    if return_Pw:
        return Sigma, np.stack(Pws)
    return Sigma
