"""CPU restatement of the Dyson step and the chemical-potential search -- TEST INFRASTRUCTURE ONLY.

Restates ``dyson<G, S1, St>`` (``src/dyson.cpp:19-236``) for one process:

* ``compute_G`` (:174-236): Sigma1 and Sigma(tau) made Hermitian (in place in the reference),
  Sigma(i w_n) = tau_to_omega(Sigma(tau)); G(i w_n) = [(i w_n + mu) S - H - Sigma1 - Sigma(i w_n)]^-1;
  G(tau) = omega_to_tau(G(i w_n)); per (spin, irreducible k).
* ``selfenergy_eigenspectra`` (:48-68): eigenvalues of S^-1 (H + Sigma1 + Sigma(i w_n)).
* ``compute_number_of_electrons`` (:70-98): N(mu) = -pref Re sum_n T_tf[nts-1, n]
  sum_{s,k} (w_k / nk) sum_i 1 / (i w_n + mu - eps_i), pref = 2 for ns = 1 without X2C, else 1;
  w_k the number of mesh points the irreducible point stands for.
* ``find_mu`` (:100-172): bracket by steps of 0.5, then bisection until |N - nel| / nel <= tol or
  the bracket is narrower than 0.01 tol; a bisection step that changes N by less than 0.001 tol
  while still off target raises (``mbpt_chemical_potential_search_failure``).

The fermionic transforms (green-grids' tau_to_omega / omega_to_tau, a dependency not in the
reference tree) are passed in as matrices T_ft [nw, nts], T_tf [nts, nw] with the Matsubara
frequencies w_n.  Only ``tests/`` imports this module.
"""
from __future__ import annotations

import numpy as np


class ChemicalPotentialSearchFailure(RuntimeError):
    """Mirror of ``mbpt_chemical_potential_search_failure``."""


def _herm(x):
    return 0.5 * (x + np.conj(np.swapaxes(x, -1, -2)))


def compute_G(H, S, Sigma1, Sigma_tau, mu, T_ft, T_tf, wn):
    """G(tau) [nts, ns, ink, nso, nso]; H, S, Sigma1 [ns, ink, nso, nso]."""
    S1 = _herm(np.asarray(Sigma1, np.complex128))
    St = _herm(np.asarray(Sigma_tau, np.complex128))
    Sw = np.tensordot(T_ft, St, axes=(1, 0))                       # [nw, ns, ink, nso, nso]
    z = 1j * np.asarray(wn, float) + mu
    M = z[:, None, None, None, None] * S[None] - H[None] - S1[None] - Sw
    Gw = np.linalg.inv(M)
    return np.tensordot(T_tf, Gw, axes=(1, 0))


def eigenspectra(H, S, Sigma1, Sigma_tau, T_ft):
    """[nw, ns, ink, nso] eigenvalues of S^-1 (H + Sigma1 + Sigma(i w_n))."""
    S1 = _herm(np.asarray(Sigma1, np.complex128))
    St = _herm(np.asarray(Sigma_tau, np.complex128))
    Sw = np.tensordot(T_ft, St, axes=(1, 0))
    X = H[None] + S1[None] + Sw
    return np.linalg.eigvals(np.linalg.solve(S[None], X))


def number_of_electrons(mu, eigs, wn, T_tf, weights, nk, ns, X2C=False):
    """``compute_number_of_electrons``; weights [ink] (mesh points per irreducible point)."""
    z = 1j * np.asarray(wn, float) + mu
    tr = np.sum((np.asarray(weights, float) / nk)[None, None, :, None] / (z[:, None, None, None] - eigs),
                axis=(1, 2, 3))
    pref = 1 if (ns == 2 or X2C) else 2
    return float(pref * -(T_tf[-1] @ tr).real)


def find_mu(mu0, nel_target, tol, N):
    """``find_mu`` on a callable N(mu); returns (N at the result, mu)."""
    mu, delta = mu0, 0.5
    nel = N(mu)
    if abs((nel - nel_target) / nel_target) > tol:
        if nel > nel_target:
            mu1, d = mu - delta, delta
            while True:
                nel1 = N(mu1)
                mu1 -= d
                if not nel1 > nel_target:
                    break
            mu2, nel2 = mu, nel
        else:
            mu2, d = mu + delta, delta
            while True:
                nel2 = N(mu2)
                mu2 += d
                if not nel2 < nel_target:
                    break
            mu1, nel1 = mu, nel
        while abs((nel - nel_target) / nel_target) > tol and abs(mu2 - mu1) > 0.01 * tol:
            mu = 0.5 * (mu1 + mu2)
            nel_old = N(mu)
            if abs((nel_old - nel_target) / nel_target) > tol and abs(nel - nel_old) < 0.001 * tol:
                raise ChemicalPotentialSearchFailure("Chemical potential search failed.")
            nel = nel_old
            if nel > nel_target:
                mu2 = mu
            else:
                mu1 = mu
    return nel, mu
