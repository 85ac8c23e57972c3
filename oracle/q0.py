"""CPU restatement of the GW q -> 0 (EXTRAPOLATE) finite-size correction -- TEST INFRASTRUCTURE ONLY.

Restates ``mbpt_q0_utils_t`` (``src/mbpt_q0_utils_t.cpp:11-178``, ``src/green/mbpt/mbpt_q0_utils_t.h``)
and its two call sites in ``gw_cpu_kernel`` (``src/gw_cpu_kernel.cpp:59-64,266-270``):

* per irreducible q (row iq of AqQ, q_ir its full-mesh index), after the Dyson step:
  eps(n, iq) = conj(A_iq)^T P-tilde(i Omega_n, q_ir) A_iq                      (aux_to_PW_00)
* per frequency n: fit Re eps(n, .) over the distinct |q| in (0, 1] (first occurrence of each
  value) with a degree-2 polynomial by the normal equations; the constant term is eps_q0(n)
  (extrapolate_q0, filter_q_abs, polyfit)
* eps_q0(tau) = w_b_to_tau_f(eps_q0(Omega)); then on the irreducible k points
  Sigma(t, s, k) -= madelung eps_q0(t) S(s, k) G(t, s, k) S(s, k)              (apply_q0_correction)
  and for X2C the aa, bb and ab blocks with S_aa, ba = ab^H                    (apply_q0_correction_2C)

madelung, AqQ [inq, NQ] and q_abs [inq] come from ``<df_path>/AqQ.h5``.  Only ``tests/`` imports it.
"""
from __future__ import annotations

import numpy as np


def eps_inv_wq(Pw_q: np.ndarray, AqQ: np.ndarray, q_ir) -> np.ndarray:
    """[nw_b, inq]; Pw_q [nq, nw_b, NQ, NQ] (full-mesh q), q_ir [inq]."""
    inq = AqQ.shape[0]
    nw = Pw_q.shape[1]
    out = np.zeros((nw, inq), np.complex128)
    for iq in range(inq):
        a = AqQ[iq]
        out[:, iq] = np.einsum("p,npq,q->n", np.conj(a), Pw_q[int(q_ir[iq])], a)
    return out


def filter_q_abs(q_abs, q_max: float = 1.0):
    """Indices of the first occurrence of each distinct |q| in (0, q_max] (exact comparison)."""
    seen, idx = set(), []
    for i, v in enumerate(q_abs):
        v = float(v)
        if 0.0 < v <= q_max and v not in seen:
            seen.add(v)
            idx.append(i)
    return idx


def polyfit(x, y, order: int) -> np.ndarray:
    """Least squares through the normal equations A^T A c = A^T y (``polyfit``)."""
    A = np.vander(np.asarray(x, float), order + 1, increasing=True)
    return np.linalg.solve(A.T @ A, A.T @ np.asarray(y, float))


def extrapolate_q0(X: np.ndarray, q_abs, order: int = 2, q_max: float = 1.0) -> float:
    idx = filter_q_abs(q_abs, q_max)
    return float(polyfit([q_abs[i] for i in idx], [X[i].real for i in idx], order)[0])


def eps_q0_tau(eps_wq: np.ndarray, q_abs, T_tb: np.ndarray) -> np.ndarray:
    """eps_q0 on the fermionic tau grid [nts] (GW_q0_correction, fit order 2, q_max 1)."""
    w = np.array([extrapolate_q0(eps_wq[n], q_abs) for n in range(eps_wq.shape[0])], np.complex128)
    return T_tb @ w


def apply_q0_correction(Sigma, G, S, eps_t, madelung, X2C=False):
    """In place on Sigma [nts, ns, ink, nso, nso] with G of the same shape and S [ns, ink, nso, nso]."""
    nts, ns, ink, nso, _ = Sigma.shape
    if not X2C:
        for t in range(nts):
            for s in range(ns):
                for k in range(ink):
                    Sigma[t, s, k] -= madelung * eps_t[t] * S[s, k] @ G[t, s, k] @ S[s, k]
        return Sigma
    nao = nso // 2
    a, b = slice(0, nao), slice(nao, nso)
    for t in range(nts):
        for k in range(ink):
            Sa = S[0, k][a, a]
            c = madelung * eps_t[t]
            Sigma[t, 0, k][a, a] -= c * Sa @ G[t, 0, k][a, a] @ Sa
            Sigma[t, 0, k][b, b] -= c * Sa @ G[t, 0, k][b, b] @ Sa
            off = c * Sa @ G[t, 0, k][a, b] @ Sa
            Sigma[t, 0, k][a, b] -= off
            Sigma[t, 0, k][b, a] -= np.conj(off.T)
    return Sigma
