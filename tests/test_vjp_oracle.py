"""Pins of the reference gradients (oracle/gw_vjp.py) -- CPU only.

* the dense torch forward equals the loop-nest oracle (gw_oracle.gw_sigma);
* the hand-derived VJP (the algorithm the device runs) equals torch autograd of the dense
  forward, for ns = 1, ns = 2 and X2C spin blocks;
* both equal central finite differences of the loop-nest oracle along random directions
  (step 1e-6, the step the device check uses);
* a VJP with one adjoint term dropped (planted in the oracle) is rejected by the same checks.
"""
import numpy as np
import pytest

from helpers import rel_err, synthetic
from oracle import gw_oracle as go
from oracle import gw_vjp as gv

CASES = [
    dict(nk=2, mesh=(2, 1, 1), nao=3, NQ=5, ns=1),
    dict(nk=3, mesh=(3, 1, 1), nao=2, NQ=4, ns=2),
    dict(nk=2, mesh=(2, 1, 1), nao=2, NQ=4, ns=1, nso=4),
]


def _system(case, seed=4):
    nso = case.get("nso", case["nao"])
    G, V, T, d = synthetic(case["nk"], case["nao"], case["NQ"], ns=case["ns"], beta=20.0, lam=1e3,
                           mesh=case["mesh"], seed=seed, nso=nso, sigma=0.3)
    rng = np.random.default_rng(seed + 100)
    Sbar = rng.standard_normal(G.shape) + 1j * rng.standard_normal(G.shape)
    Sbar = 0.5 * (Sbar + np.conj(np.swapaxes(Sbar, -1, -2)))
    args = (V, T.T_bt, T.T_tb, d["nk"], d["nao"], d["nso"], d["ns"], d["NQ"], d["mesh"])
    return G, Sbar, args, d


def _fd(G, Sbar, dG, args, h=1e-6):
    V, T_bt, T_tb, nk, nao, nso, ns, NQ, mesh = args
    Sp = go.gw_sigma(G + h * dG, V, T_bt, T_tb, nk, nao, nso, ns, NQ, mesh=mesh)
    Sm = go.gw_sigma(G - h * dG, V, T_bt, T_tb, nk, nao, nso, ns, NQ, mesh=mesh)
    return float(np.real(np.vdot(Sbar, (Sp - Sm) / (2 * h))))


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_dense_forward_matches_oracle(ci):
    G, _, args, d = _system(CASES[ci])
    V, T_bt, T_tb, nk, nao, nso, ns, NQ, mesh = args
    import torch
    S_dense = gv.sigma_dense(torch.tensor(G), gv.v_full(V, nk), T_bt, T_tb, nk, nao, nso, ns, NQ, mesh).numpy()
    S_loop = go.gw_sigma(G, V, T_bt, T_tb, nk, nao, nso, ns, NQ, mesh=mesh)
    assert rel_err(S_dense, S_loop) < 1e-13


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_manual_vjp_matches_autograd_and_fd(ci):
    G, Sbar, args, d = _system(CASES[ci])
    _, g_ag = gv.vjp_autograd(G, Sbar, *args)
    g_m = gv.vjp_manual(G, Sbar, *args)
    assert rel_err(g_m, g_ag) < 1e-12
    rng = np.random.default_rng(7)
    for _ in range(2):
        dG = rng.standard_normal(G.shape) + 1j * rng.standard_normal(G.shape)
        fd = _fd(G, Sbar, dG, args)
        an = float(np.real(np.vdot(g_m, dG)))
        assert abs(fd - an) < 1e-7 * max(abs(an), 1.0)


@pytest.mark.parametrize("plant", ["no_inverse_adjoint", "no_gbeta_term"])
def test_planted_mistakes_are_rejected(plant):
    """A VJP missing one adjoint term fails against autograd and against finite differences."""
    G, Sbar, args, d = _system(CASES[0])
    _, g_ag = gv.vjp_autograd(G, Sbar, *args)
    g_bad = gv.vjp_manual(G, Sbar, *args, plant=plant)
    assert rel_err(g_bad, g_ag) > 1e-6
    dG = np.random.default_rng(3).standard_normal(G.shape) + 0j
    fd = _fd(G, Sbar, dG, args)
    assert abs(fd - float(np.real(np.vdot(g_bad, dG)))) > 1e-6 * abs(fd)
