"""Pins of the CPU oracle (oracle/gw_oracle.py) -- CPU only.

The oracle is trusted only through these checks:
* the reference's own GW fixtures (test/solvers_test.cpp:391 GW_X2C and :602-672 SEET DC),
  at the reference's tolerance 1e-6 abs and at the level actually reached;
* closed forms (scalar nao = NQ = nk = 1 system);
* an independent brute-force einsum of the whole chain on tiny random systems, written
  over full spinor indices rather than the reference's spin-block loops;
* planted mistakes (in the oracle, on the CPU) that the fixtures must reject;
* invariants (Hermiticity, tau symmetry, zero-padded aux rows).
"""
import numpy as np
import pytest

from helpers import load_fixture, rel_err, synthetic, transformer
from oracle import gw_oracle as go


def _run_fixture(name, **kw):
    G, V, T, d, S_ref = load_fixture(name)
    S = go.gw_sigma(G, V, T.T_bt, T.T_tb, d["nk"], d["nao"], d["nso"], d["ns"], d["NQ"], **kw)
    return S, S_ref


def test_fixture_gw_x2c():
    """GW_X2C section: nk=1, X2C nao=10 nso=20, NQ=154, IR 1e4, beta=100."""
    S, S_ref = _run_fixture("gw_x2c")
    err = np.abs(S - S_ref).max()
    assert err < 1e-6          # reference tolerance (solvers_test.cpp:134)
    assert err < 1e-9          # reached: ~1.2e-11 with this project's IR basis


def test_fixture_seet_dc():
    """SEET DC TEST inputs: nk=1, ns=2, NQ=36 with 3 stored aux rows."""
    S, S_ref = _run_fixture("seet_dc")
    assert np.abs(S - S_ref).max() < 1e-6


def test_zero_padded_aux_rows_do_not_change_sigma():
    G, V, T, d, _ = load_fixture("seet_dc")
    S36 = go.gw_sigma(G, V, T.T_bt, T.T_tb, 1, d["nao"], d["nso"], d["ns"], 36)
    S3 = go.gw_sigma(G, V[:, :3].copy(), T.T_bt, T.T_tb, 1, d["nao"], d["nso"], d["ns"], 3)
    assert np.abs(S36 - S3).max() < 1e-14


def _plant(monkeypatch, mistake):
    if mistake == "prefactor":
        orig = go.P0_contraction
        monkeypatch.setattr(go, "P0_contraction", lambda Gb, Gq, vm, vmm, P0t, pref, nao, NQ:
                            orig(Gb, Gq, vm, vmm, P0t, 2.0 * pref, nao, NQ))
    elif mistake == "no_conj":
        def bad(Gb, Gq, vm, vmm, P0t, pref, nao, NQ):
            X1m = Gb @ vm.T.reshape(nao, nao * NQ)
            X2m = vmm @ Gq.T                      # conj dropped
            P0t.T[...] -= pref * (X2m.reshape(NQ, nao * nao) @ X1m.reshape(nao * nao, NQ))
        monkeypatch.setattr(go, "P0_contraction", bad)
    elif mistake == "transpose_G":
        orig = go.selfenergy_contraction
        monkeypatch.setattr(go, "selfenergy_contraction", lambda Gk, vm, P, nao, NQ:
                            orig(Gk.T.copy(), vm, P, nao, NQ))
    elif mistake == "no_mirror":
        def bad_sym(P0, nk):
            nts = P0.shape[0]
            for it in range(nts // 2):
                P0[it] /= nk
                P0[it] = go.make_hermitian(P0[it])
        monkeypatch.setattr(go, "symmetrize_P0", bad_sym)
    elif mistake == "k_plus_for_minus":
        orig = go.k_add
        monkeypatch.setattr(go, "k_add", lambda k, q, mesh, sign=1: orig(k, q, mesh, +1))
    elif mistake == "no_adjoint_for_k1_lt_k2":
        monkeypatch.setattr(go, "v_pair", lambda Vp, k1, k2: Vp[go.pair_index(k1, k2)])


@pytest.mark.parametrize("mistake", ["prefactor", "transpose_G", "no_mirror"])
def test_planted_mistakes_fail_fixture(monkeypatch, mistake):
    """Slips the reference's Sigma fixture rejects (its V is real, nk = 1)."""
    _plant(monkeypatch, mistake)
    S, S_ref = _run_fixture("gw_x2c")
    assert np.abs(S - S_ref).max() > 1e-6


@pytest.mark.parametrize("mistake", ["prefactor", "no_conj", "transpose_G", "no_mirror",
                                     "k_plus_for_minus", "no_adjoint_for_k1_lt_k2"])
def test_planted_mistakes_fail_brute_force(monkeypatch, mistake):
    """Slips the complex, multi-k brute-force einsum rejects (conjugation, k maps, pair adjoint)."""
    mesh = (3, 1, 1)
    G, V, T, d = synthetic(3, 3, 4, beta=20.0, lam=1e3, mesh=mesh, seed=5, sigma=0.3)
    S_bf = _brute_force(G, V, T, 3, 3, 3, 1, 4, mesh)
    _plant(monkeypatch, mistake)
    S = go.gw_sigma(G, V, T.T_bt, T.T_tb, 3, 3, 3, 1, 4, mesh=mesh)
    assert rel_err(S, S_bf) > 1e-6


def test_scalar_closed_form():
    """nao = NQ = nk = 1: P0(tau) = -2|v|^2 G(beta-tau) G(tau); P~(iW) = P0/(1-P0); Sigma = -|v|^2 G P~."""
    beta, lam = 50.0, 1e3
    T = transformer(lam, beta)
    eps, v = 0.37, 0.8 + 0.3j
    f = np.where(eps >= 0, np.exp(-eps * T.tau) / (1 + np.exp(-beta * eps)),
                 np.exp(eps * (beta - T.tau)) / (1 + np.exp(beta * eps)))
    G = (-f).astype(complex).reshape(-1, 1, 1, 1, 1)
    S, Pt = go.gw_sigma(G, np.array([[[[v]]]]), T.T_bt, T.T_tb, 1, 1, 1, 1, 1, return_P=True)
    g = G[:, 0, 0, 0, 0]
    P0 = -2.0 * abs(v) ** 2 * g[::-1] * g
    P0w = T.T_bt @ P0
    Pt_ref = T.T_tb @ (P0w / (1 - P0w))
    assert np.abs(Pt[0, :, 0, 0] - Pt_ref).max() < 1e-13
    assert np.abs(S[:, 0, 0, 0, 0] - (-abs(v) ** 2 * g * Pt_ref)).max() < 1e-13


def _brute_force(G, Vp, T, nk, nao, nso, ns, NQ, mesh):
    """Whole chain with einsum over full spinor indices (independent of the block loops)."""
    nts = G.shape[0]
    X2C = nso != nao
    pref = 1.0 if (ns == 2 or X2C) else 2.0
    pns = nso // nao
    Gr = G.reshape(nts, ns, nk, pns, nao, pns, nao)
    # full V(k1, k2) from the packed k1 >= k2 store, and k +/- q, written independently
    Vfull = np.empty((nk, nk, NQ, nao, nao), complex)
    for k1 in range(nk):
        for k2 in range(k1 + 1):
            blk = Vp[k1 * (k1 + 1) // 2 + k2]
            Vfull[k1, k2] = blk
            Vfull[k2, k1] = np.conj(blk).transpose(0, 2, 1)
    coords = np.array(np.unravel_index(np.arange(nk), mesh)).T

    def shift(k, q, sign):
        return int(np.ravel_multi_index(tuple((coords[k] + sign * coords[q]) % np.array(mesh)), mesh))

    S = np.zeros_like(G)
    for q in range(nk):
        P0 = np.zeros((nts, NQ, NQ), complex)
        for k in range(nk):
            kq = shift(k, q, +1)
            v = Vfull[k, kq]
            for t in range(nts // 2):
                for s in range(ns):
                    Gb = Gr[nts - 1 - t, s, k]       # [b, t_, a, p]
                    Gq = Gr[t, s, kq]                # [a, m, b, n]
                    P0[t] -= pref * np.einsum("btap,Qpm,ambn,Rtn->QR", Gb, v, Gq, np.conj(v))
        P0 /= nk
        for t in range(nts // 2):
            P0[t] = 0.5 * (P0[t] + P0[t].conj().T)
            P0[nts - 1 - t] = P0[t]
        Pw = np.einsum("nt,tQR->nQR", T.T_bt, P0)
        I = np.eye(NQ)
        Ptw = np.stack([np.linalg.solve(I - Pw[n], Pw[n]) for n in range(Pw.shape[0])])
        Pt = np.einsum("tn,nQR->tQR", T.T_tb, Ptw)
        for k in range(nk):
            kmq = shift(k, q, -1)
            v = Vfull[k, kmq]
            for t in range(nts):
                for s in range(ns):
                    blk = np.einsum("Qim,ambn,QR,Rjn->aibj", v, Gr[t, s, kmq], Pt[t], np.conj(v))
                    S[t, s, k] -= blk.reshape(nso, nso)
    return S / nk


@pytest.mark.parametrize("case", [
    dict(nk=2, mesh=(2, 1, 1), nao=3, NQ=5, ns=1),
    dict(nk=4, mesh=(2, 2, 1), nao=2, NQ=4, ns=2),
    dict(nk=2, mesh=(2, 1, 1), nao=2, NQ=3, ns=1, nso=4),      # X2C spin blocks
])
def test_brute_force_einsum(case):
    nso = case.get("nso", case["nao"])
    G, V, T, d = synthetic(case["nk"], case["nao"], case["NQ"], ns=case["ns"], beta=20.0, lam=1e3,
                           mesh=case["mesh"], seed=3, nso=nso, sigma=0.3)
    S = go.gw_sigma(G, V, T.T_bt, T.T_tb, d["nk"], d["nao"], nso, d["ns"], d["NQ"], mesh=case["mesh"])
    S_bf = _brute_force(G, V, T, d["nk"], d["nao"], nso, d["ns"], d["NQ"], case["mesh"])
    assert rel_err(S, S_bf) < 1e-12


def test_invariants_config_small():
    G, V, T, d = synthetic(8, 4, 10, beta=50.0, lam=1e3, mesh=(2, 2, 2), seed=1)
    S, Pt = go.gw_sigma(G, V, T.T_bt, T.T_tb, 8, 4, 4, 1, 10, mesh=(2, 2, 2), return_P=True)
    # Sigma Hermitian per (tau, s, k); P~ Hermitian and symmetric under tau -> beta - tau
    assert np.abs(S - np.conj(np.swapaxes(S, -1, -2))).max() < 1e-12 * np.abs(S).max()
    assert np.abs(Pt - np.conj(np.swapaxes(Pt, -1, -2))).max() < 1e-10 * np.abs(Pt).max()
    assert np.abs(Pt - Pt[:, ::-1]).max() < 1e-9 * np.abs(Pt).max()
    # P~(tau) is negative semidefinite at tau = 0 (screening): diag real and <= 0
    assert np.all(np.real(np.einsum("qii->qi", Pt[:, 0])) <= 1e-12)


def test_odd_nts_raises_wrong_grid():
    G = np.zeros((5, 1, 1, 2, 2), complex)
    with pytest.raises(go.WrongGrid):
        go.gw_sigma(G, np.zeros((1, 3, 2, 2), complex), np.zeros((4, 5)), np.zeros((5, 4)), 1, 2, 2, 1, 3)


def test_gm_energy_closed_form():
    """GM energy of scalar single poles: E = -pref Re int_0^beta G(beta - tau) Sigma(tau) dtau."""
    beta, lam = 50.0, 1e3
    T = transformer(lam, beta)
    a, b, c = 0.4, -0.25, 0.05

    def gpole(e, tau):
        return np.where(e >= 0, -np.exp(-e * tau) / (1 + np.exp(-beta * e)),
                        -np.exp(e * (beta - tau)) / (1 + np.exp(beta * e)))
    G = gpole(a, T.tau).astype(complex).reshape(-1, 1, 1, 1, 1)
    S = (c * gpole(b, T.tau)).astype(complex).reshape(-1, 1, 1, 1, 1)
    # int_0^beta G_a(beta - tau) G_b(tau) dtau for single poles
    na, nb = 1 / (1 + np.exp(-beta * a)), 1 / (1 + np.exp(-beta * b))
    integral = na * nb * np.exp(-a * beta) * (np.exp((a - b) * beta) - 1) / (a - b)
    E = go.gm_energy(G, S, T.T_ft, T.T_tf, nk=1, ns=1)
    assert abs(E - (-1.0 * c * integral)) < 1e-10
