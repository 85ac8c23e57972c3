"""CPU FP64 oracle for the GW self-energy hot path -- TEST INFRASTRUCTURE ONLY.

This module is the checker, not the product: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import it.  The
GPU library never calls it, and the product path fails when the HIP library is
missing instead of falling back here.

It restates ``green::mbpt::kernels::gw_cpu_kernel::solve`` in numpy (complex128,
zgemm through numpy's BLAS), keeping the reference's loop nest and GEMM order:

* ``gw_sigma``            -- ``src/gw_cpu_kernel.cpp:31-72`` (solve: zero Sigma, loop q,
                             Sigma /= nk); full k mesh, one process, so the q loop is
                             every q and the inter-node allreduce is the identity.
* ``selfenergy_innerloop`` -- ``:74-134`` (odd nts -> error; P0 over k1; symmetrize;
                             P-tilde; Sigma over k1 and the star of q).
* ``eval_P0_tilde``       -- ``:143-190`` (half tau grid, spin blocks, prefactor).
* ``P0_contraction``      -- ``:200-215`` (three GEMMs, P0^T -= pref * X2 X1).
* ``symmetrize_P0``       -- ``:217-225`` (/nk, Hermitise, P0(beta - tau) = P0(tau)).
* ``eval_P_tilde``        -- ``:227-279`` (tau_f -> w_b, (I - P0)^-1 Hermitised times P0,
                             w_b -> tau_f).  The reference inverts with Eigen's LDLT;
                             here LAPACK zgesv solves against the identity -- the same
                             inverse up to rounding.
* ``eval_selfenergy``     -- ``:295-354`` and ``selfenergy_contraction`` ``:359-373``.
* ``v_pair``              -- ``df_integral_t::symmetrize`` (``df_integral_t.h:162-187``) with
                             the pair index of ``read_integrals`` (``:64-84``) for a mesh
                             without symmetry (no conj/trans redirection).
* ``gm_energy``           -- ``compute_energy`` (``src/common_utils.cpp:5-55``),
                             correlation part (Galitskii-Migdal); ``compute_energy`` all three
                             energies (one-body, Hartree-Fock, correlation).

k/q maps follow the variable names at ``gw_cpu_kernel.cpp:90,117,311``: k1_from_k2q(k, q)
= k + q and k2_from_k1q(k, q) = k - q on a Monkhorst-Pack mesh, componentwise modulo the
mesh size, with k = (i0 * n1 + i1) * n2 + i2.

Precision switches mirror the reference's ``P_sp`` / ``Sigma_sp`` (``common_defs.h:162-163``):
products in complex64, cast to complex128 before accumulation.
"""
from __future__ import annotations

import numpy as np


class WrongGrid(RuntimeError):
    """Mirror of ``mbpt_wrong_grid`` (``src/green/mbpt/except.h:33-36``)."""


# ---------------------------------------------------------------- k-mesh maps
def mesh_shape(nk: int, mesh=None):
    if mesh is not None:
        m = tuple(int(v) for v in mesh)
        assert int(np.prod(m)) == nk
        return m
    n = round(nk ** (1.0 / 3.0))
    if n ** 3 == nk:
        return (n, n, n)
    return (nk, 1, 1)


def k_add(k: int, q: int, mesh, sign: int = 1) -> int:
    n0, n1, n2 = mesh
    a = np.unravel_index(k, mesh)
    b = np.unravel_index(q, mesh)
    c = [(int(a[i]) + sign * int(b[i])) % mesh[i] for i in range(3)]
    return int(np.ravel_multi_index(c, mesh))


def pair_index(k1: int, k2: int) -> int:
    """``df_integral_t.h:68``: idx = k1 (k1 + 1)/2 + k2 for k1 >= k2, else swapped."""
    return k1 * (k1 + 1) // 2 + k2 if k1 >= k2 else k2 * (k2 + 1) // 2 + k1


def v_pair(Vp: np.ndarray, k1: int, k2: int) -> np.ndarray:
    """V^Q_ij(k1, k2) from the packed k1 >= k2 store (``df_integral_t.h:162-177``).

    Vp: [npair, NQ, nao, nao]; for k1 < k2 each Q block is the adjoint of the stored
    (k2, k1) block.
    """
    v = Vp[pair_index(k1, k2)]
    if k1 < k2:
        return np.ascontiguousarray(np.conj(np.transpose(v, (0, 2, 1))))
    return v


def make_hermitian(x: np.ndarray) -> np.ndarray:
    """``common_defs.h:104-115``: X = (X + X^H)/2 on the two innermost axes."""
    return 0.5 * (x + np.conj(np.swapaxes(x, -1, -2)))


# ---------------------------------------------------------------- P0
def eval_P0_tilde(P0, G, Vp, q, nk, nao, nso, ns, NQ, X2C, mesh, P_sp=False):
    """Accumulate P0(tau, q) over k1 for the half tau grid (``gw_cpu_kernel.cpp:89-102,143-190``)."""
    nts = G.shape[0]
    prec = np.complex64 if P_sp else np.complex128
    pns = nso // nao
    pref = 1.0 if (ns == 2 or X2C) else 2.0
    pseudo_ns = ns if not X2C else 4
    half = nts // 2
    for k1 in range(nk):
        k1q = k_add(k1, q, mesh, +1)
        v = v_pair(Vp, k1, k1q).astype(prec)                 # (Q, p, m)
        vm = v.reshape(NQ, nao * nao)
        vmm = v.reshape(NQ * nao, nao)
        for t in range(half):
            tt = nts - t - 1
            for s in range(pseudo_ns):
                is_ = s // (pns * pns)
                a = (s // pns) % pns
                b = s % pns
                Gb = G[tt, is_, k1][b * nao:(b + 1) * nao, a * nao:(a + 1) * nao].astype(prec)
                Gq = G[t, is_, k1q][a * nao:(a + 1) * nao, b * nao:(b + 1) * nao].astype(prec)
                P0_contraction(Gb, Gq, vm, vmm, P0[t], pref, nao, NQ)


def P0_contraction(Gb, Gq, vm, vmm, P0t, pref, nao, NQ):
    """``gw_cpu_kernel.cpp:200-215``."""
    VVm = vm.T                                               # (pm, Q)
    VVmm = VVm.reshape(nao, nao * NQ)                        # (p, mQ) view of the same buffer
    X1m = Gb @ VVmm                                          # (t, mQ)
    X2m = np.conj(vmm) @ Gq.T                                # (Q't, m)
    X2mm = X2m.reshape(NQ, nao * nao)
    X1mm = X1m.reshape(nao * nao, NQ)
    P0t.T[...] -= pref * (X2mm @ X1mm).astype(np.complex128)


def symmetrize_P0(P0, nk):
    """``gw_cpu_kernel.cpp:217-225``."""
    nts = P0.shape[0]
    for it in range(nts // 2):
        P0[it] /= nk
        P0[it] = make_hermitian(P0[it])
        P0[nts - it - 1] = P0[it]


def eval_P_tilde(P0, T_bt, T_tb, return_w=False):
    """``gw_cpu_kernel.cpp:227-279`` on one q: returns P-tilde(tau) [nts, NQ, NQ] (and P-tilde on the
    bosonic frequencies [nw_b, NQ, NQ] when ``return_w``, the array ``aux_to_PW_00`` reads at :267-270)."""
    NQ = P0.shape[-1]
    Pw = np.tensordot(T_bt, P0, axes=(1, 0))                 # tau_f_to_w_b
    I = np.eye(NQ, dtype=np.complex128)
    for n in range(Pw.shape[0]):
        temp = I - Pw[n]
        temp = np.linalg.solve(temp, I)
        temp = 0.5 * (temp + np.conj(temp.T))
        Pw[n] = temp @ Pw[n]
    Pt = np.tensordot(T_tb, Pw, axes=(1, 0))                 # w_b_to_tau_f
    return (Pt, Pw) if return_w else Pt


# ---------------------------------------------------------------- Sigma
def eval_selfenergy(Sigma, G, Vp, Pt, q, nk, nao, nso, ns, NQ, X2C, mesh, Sigma_sp=False):
    """Accumulate -V G P V^H for every k1 at this q (``gw_cpu_kernel.cpp:113-130,295-373``)."""
    nts = G.shape[0]
    prec = np.complex64 if Sigma_sp else np.complex128
    pns = nso // nao
    pseudo_ns = ns if not X2C else 4
    Pp = Pt.astype(prec)
    for k1 in range(nk):
        k1mq = k_add(k1, q, mesh, -1)
        v = v_pair(Vp, k1, k1mq).astype(prec)                 # (Q, i, m)
        vm = v.reshape(NQ * nao, nao)
        for t in range(nts):
            P = Pp[t]
            for s in range(pseudo_ns):
                is_ = s // (pns * pns)
                a = (s // pns) % pns
                b = s % pns
                Gk = G[t, is_, k1mq][a * nao:(a + 1) * nao, b * nao:(b + 1) * nao].astype(prec)
                Sm = selfenergy_contraction(Gk, vm, P, nao, NQ)
                Sigma[t, is_, k1][a * nao:(a + 1) * nao, b * nao:(b + 1) * nao] -= Sm


def selfenergy_contraction(Gk, vm, P, nao, NQ):
    """``gw_cpu_kernel.cpp:359-373``."""
    Y1m = vm @ Gk                                            # (Qi, n)
    Y1mm = Y1m.reshape(NQ, nao * nao)
    Y2mm = Y1mm.T @ P                                        # (in, Q')
    X2m = np.conj(vm.T)                                      # (n, Q'j)
    X2mm = X2m.reshape(nao * NQ, nao)                        # (nQ', j)
    Y2mmm = Y2mm.reshape(nao, nao * NQ)                      # (i, nQ')
    return (Y2mmm @ X2mm).astype(np.complex128)


# ---------------------------------------------------------------- driver
def gw_sigma(G, Vp, T_bt, T_tb, nk, nao, nso, ns, NQ, X2C=None, mesh=None,
             P_sp=False, Sigma_sp=False, return_P=False, return_Pw=False):
    """Sigma(tau, s, k) from G(tau, s, k) -- ``gw_cpu_kernel::solve``.

    G: [nts, ns, nk, nso, nso] complex128; Vp: [nk(nk+1)/2, NQ, nao, nao] complex128;
    T_bt: [nw_b, nts]; T_tb: [nts, nw_b].  Returns Sigma with G's shape (and the
    P-tilde(tau, q) stack [nq, nts, NQ, NQ] when ``return_P``, or the frequency-domain stack
    [nq, nw_b, NQ, NQ] when ``return_Pw``).
    """
    G = np.asarray(G, dtype=np.complex128)
    nts = G.shape[0]
    if X2C is None:
        X2C = nao != nso
    if nts % 2 != 0:
        raise WrongGrid("Number of tau points should be even!!!")
    mesh = mesh_shape(nk, mesh)
    Sigma = np.zeros_like(G)
    Ps = []
    for q in range(nk):
        P0 = np.zeros((nts, NQ, NQ), dtype=np.complex128)
        eval_P0_tilde(P0, G, Vp, q, nk, nao, nso, ns, NQ, X2C, mesh, P_sp)
        symmetrize_P0(P0, nk)
        Pt, Pw = eval_P_tilde(P0, T_bt, T_tb, return_w=True)
        if return_P:
            Ps.append(Pt)
        elif return_Pw:
            Ps.append(Pw)
        eval_selfenergy(Sigma, G, Vp, Pt, q, nk, nao, nso, ns, NQ, X2C, mesh, Sigma_sp)
    Sigma /= nk
    if return_P or return_Pw:
        return Sigma, np.stack(Ps)
    return Sigma


def compute_energy(G, Sigma1, Sigma, H, T_ft, T_tf, nk, ns, X2C=False, weights=None):
    """``compute_energy`` (``src/common_utils.cpp:5-55``): (E_1b, E_HF, E_corr).

    G, Sigma [nts, ns, ink, nso, nso]; Sigma1, H [ns, ink, nso, nso]; weights [ink] (mesh points per
    irreducible point; None: one each); nkpw = 1/nk.  dm = -(2 or 1) G(beta) (``:16-17``);
    E_1b = nkpw sum_{s,k} w_k Re tr(dm H) (``:50``), E_HF = nkpw sum 1/2 w_k Re tr(dm (Sigma1 + 2 H))
    (``:47-48``), E_corr = -pref Re[T_tf[beta] . sum_{s,k} w_k tr(G(i w) Sigma(i w))] nkpw (``:28-41``).
    """
    G = np.asarray(G)
    ink = G.shape[2]
    w = np.ones(ink) if weights is None else np.asarray(weights, dtype=float)
    c = 1.0 if (ns == 2 or X2C) else 2.0
    dm = -c * G[-1]
    e1 = 0.0
    ehf = 0.0
    for s in range(ns):
        for k in range(ink):
            e1 += w[k] * np.trace(dm[s, k] @ H[s, k]).real
            ehf += 0.5 * w[k] * np.trace(dm[s, k] @ (Sigma1[s, k] + 2.0 * H[s, k])).real
    Gw = np.tensordot(T_ft, G, axes=(1, 0))
    Sw = np.tensordot(T_ft, Sigma, axes=(1, 0))
    GS_w = np.einsum("k,nskij,nskji->n", w, Gw, Sw)
    pref = 1.0 if (ns == 1 and not X2C) else 0.5
    ecorr = -(T_tf[-1] @ GS_w).real * pref
    return e1 / nk, ehf / nk, float(ecorr / nk)


def gm_energy(G, Sigma, T_ft, T_tf, nk, ns, X2C=False):
    """Galitskii-Migdal correlation energy, ``compute_energy`` (``src/common_utils.cpp:5-55``).

    E_corr = -pref * Re[ sum_n T_tf[beta, n] * sum_{s,k} w_k Tr(G(i w_n) Sigma(i w_n)) ] * nkpw,
    pref = 1 for ns = 1 without X2C, else 1/2; full mesh: w_k = 1, nkpw = 1/nk.
    """
    Gw = np.tensordot(T_ft, G, axes=(1, 0))
    Sw = np.tensordot(T_ft, Sigma, axes=(1, 0))
    GS_w = np.einsum("nskij,nskji->n", Gw, Sw)
    GS_beta = T_tf[-1] @ GS_w
    pref = 1.0 if (ns == 1 and not X2C) else 0.5
    return float(-(GS_beta.real) * pref / nk)
