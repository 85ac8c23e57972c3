"""Shared fixtures-as-functions for the tests (test infrastructure)."""
import os

import numpy as np

from gwhip import inputs, irgrid

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# (nk, mesh, nao, NQ, ns, beta, lam) -- the CPU-runnable workload: 2x2x2 mesh, nao 8, naux 40
CONFIG0 = dict(nk=8, mesh=(2, 2, 2), nao=8, NQ=40, ns=1, beta=50.0, lam=1e3)


def transformer(lam: float, beta: float, tau=None):
    g = irgrid.load_grid(lam)
    if tau is None:
        return irgrid.Transformer.from_grid(g, beta)
    return irgrid.Transformer.from_nodes(g, beta, tau)


def load_fixture(name: str):
    """Reference fixture -> (G, V_packed, T, dims)."""
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    nao, nso, ns, NQ, nk = (int(z[k]) for k in ("nao", "nso", "ns", "NQ", "nk"))
    if "V" in z.files:
        V = z["V"].copy()
    else:
        V = np.zeros((1, NQ, nao, nao), complex)
        V[0, :z["V_rows"].shape[1]] = z["V_rows"][0]
    T = transformer(float(z["lam"]), float(z["beta"]), z["tau"])
    dims = dict(nk=nk, nao=nao, nso=nso, ns=ns, NQ=NQ, beta=float(z["beta"]), mesh=(1, 1, 1))
    return z["G_tau"].copy(), V, T, dims, z["Sigma_tau"].copy()


def synthetic(nk, nao, NQ, ns=1, beta=50.0, lam=1e3, mesh=None, seed=0, nso=None, sigma=0.1):
    """Seeded synthetic system: (G, V_packed, T, dims)."""
    nso = nao if nso is None else nso
    T = transformer(lam, beta)
    G = inputs.generate_G(seed, nk, nso, ns, T.tau, beta)
    V = inputs.generate_V(seed, nk, NQ, nao, sigma)
    if mesh is None:
        n = round(nk ** (1.0 / 3.0))
        mesh = (n, n, n) if n ** 3 == nk else (nk, 1, 1)
    dims = dict(nk=nk, nao=nao, nso=nso, ns=ns, NQ=NQ, beta=beta, mesh=tuple(mesh))
    return G, V, T, dims


def rel_err(a, b):
    """max-abs difference over max |b|."""
    return float(np.abs(a - b).max() / np.abs(b).max())


def min_eig_I_minus_P0(G, V, T, d):
    """Smallest eigenvalue of the Hermitian part of I - P0(i Omega_n, q) over every q and bosonic
    frequency, from the oracle's P0 (the device's Dyson step reports a matrix that is not positive
    definite)."""
    from oracle import gw_oracle as go

    nk, NQ = d["nk"], d["NQ"]
    X2C = d["nso"] != d["nao"]
    lo = np.inf
    for q in range(nk):
        P0 = np.zeros((T.nts, NQ, NQ), complex)
        go.eval_P0_tilde(P0, G, V, q, nk, d["nao"], d["nso"], d["ns"], NQ, X2C, d["mesh"])
        go.symmetrize_P0(P0, nk)
        for p in np.tensordot(T.T_bt, P0, axes=(1, 0)):
            A = np.eye(NQ) - p
            lo = min(lo, np.linalg.eigvalsh(0.5 * (A + np.conj(A.T))).min())
    return float(lo)
