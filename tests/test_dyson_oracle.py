"""Pins of the Dyson / chemical-potential oracle (oracle/dyson.py) -- CPU only.

* The reference's four HF_orth_invariance fixtures (test/solvers_test.cpp:296-452): the stored
  first-iteration G(tau) is the Dyson step of the stored H, S, Sigma1, Sigma(tau) and mu; reproduced
  to ~5e-8 absolute (|G| ~ 8.5), the level at which this project's IR basis matches green-grids'.
* The stored mu (-0.25) is what find_mu returns from mu = 0 with tolerance 1e-9.
* Closed forms: with Sigma = 0 the electron count is the Fermi-function sum over the generalised
  eigenvalues of (H, S); G(beta) S traced gives the same count.
* The trace form the device uses, -pref Re sum_n T_tf[-1, n] sum (w/nk) tr(G(i w_n) S), equals the
  reference's eigenvalue form.
* Planted mistakes (in the oracle, on the CPU): sign of mu, a dropped Hermitian part, the search's
  step direction -- each rejected.
"""
import os

import numpy as np
import pytest

from helpers import GOLDEN, transformer
from oracle import dyson as od

MODES = ("none", "lowdin", "mo", "natural")


def _fixture():
    z = np.load(os.path.join(GOLDEN, "dyson_orth.npz"), allow_pickle=False)
    T = transformer(float(z["lam"]), float(z["beta"]), z["tau"])
    wn = (2 * T.n_fermi + 1) * np.pi / float(z["beta"])
    return z, T, wn


@pytest.mark.parametrize("mode", MODES)
def test_orth_fixture_G(mode):
    z, T, wn = _fixture()
    G = od.compute_G(z[mode + "_H"], z[mode + "_S"], z[mode + "_Sigma1"], z[mode + "_Sigma_tau"],
                     float(z[mode + "_mu"]), T.T_ft, T.T_tf, wn)
    err = np.abs(G - z[mode + "_G_tau"]).max()
    assert err < 1e-6      # the reference's default IsCloseTo tolerance
    assert err < 2e-7      # reached: ~5e-8 (IR basis of this project vs green-grids')


def test_orth_fixture_mu_search():
    z, T, wn = _fixture()
    e = od.eigenspectra(z["none_H"], z["none_S"], z["none_Sigma1"], z["none_Sigma_tau"], T.T_ft)
    N = lambda mu: od.number_of_electrons(mu, e, wn, T.T_tf, [1.0], 1, 2)
    nel, mu = od.find_mu(0.0, float(z["none_nel"]), 1e-9, N)
    assert mu == float(z["none_mu"]) == -0.25
    assert abs(nel - 2.0) < 1e-9


def test_count_closed_form_fermi():
    """Sigma = 0: N(mu) = pref * sum_k w_k/nk sum_i f(eps_i - mu), eps the (H, S) eigenvalues."""
    z, T, wn = _fixture()
    beta = float(z["beta"])
    H, S = z["none_H"], z["none_S"]
    zero1 = np.zeros_like(z["none_Sigma1"])
    zerot = np.zeros_like(z["none_Sigma_tau"])
    e = od.eigenspectra(H, S, zero1, zerot, T.T_ft)
    eps = np.linalg.eigvals(np.linalg.solve(S, H)).real
    for mu in (-0.3, 0.0, 0.35):
        N = od.number_of_electrons(mu, e, wn, T.T_tf, [1.0], 1, 2)
        fermi = np.sum(1.0 / (1.0 + np.exp(beta * (eps - mu))))
        assert abs(N - fermi) < 1e-7, (mu, N, fermi)


def test_trace_form_equals_eigenvalue_form():
    z, T, wn = _fixture()
    rng = np.random.default_rng(0)
    St = 0.01 * (rng.standard_normal(z["none_Sigma_tau"].shape) + 1j * rng.standard_normal(z["none_Sigma_tau"].shape))
    H, S, S1 = z["none_H"], z["none_S"], z["none_Sigma1"]
    e = od.eigenspectra(H, S, S1, St, T.T_ft)
    for mu in (-0.4, 0.1):
        N_eig = od.number_of_electrons(mu, e, wn, T.T_tf, [1.0], 1, 2)
        Sh = 0.5 * (St + np.conj(np.swapaxes(St, -1, -2)))
        Sw = np.tensordot(T.T_ft, Sh, axes=(1, 0))
        S1h = 0.5 * (S1 + np.conj(np.swapaxes(S1, -1, -2)))
        M = (1j * wn + mu)[:, None, None, None, None] * S[None] - H[None] - S1h[None] - Sw
        tr = np.einsum("wskij,skji->w", np.linalg.inv(M), S)
        N_tr = float(-(T.T_tf[-1] @ tr).real)
        assert abs(N_tr - N_eig) < 1e-10


@pytest.mark.parametrize("mistake", ["mu_sign", "no_hermitian_part"])
def test_planted_mistakes(mistake):
    z, T, wn = _fixture()
    mode = "none"
    S1 = z[mode + "_Sigma1"].copy()
    mu = float(z[mode + "_mu"])
    if mistake == "mu_sign":
        mu = -mu
    if mistake == "no_hermitian_part":   # a non-Hermitian Sigma1 used as is
        S1 = S1 + 0.05j * np.eye(S1.shape[-1])
        G = od.compute_G(z[mode + "_H"], z[mode + "_S"], S1, z[mode + "_Sigma_tau"], mu, T.T_ft, T.T_tf, wn)
        assert np.abs(G - z[mode + "_G_tau"]).max() < 1e-6     # the Hermitian part removes it ...
        Sw = 0.0
        M = (1j * wn + mu)[:, None, None, None, None] * z[mode + "_S"][None] - z[mode + "_H"][None] - S1[None] - Sw
        Gb = np.tensordot(T.T_tf, np.linalg.inv(M), axes=(1, 0))
        assert np.abs(Gb - z[mode + "_G_tau"]).max() > 1e-6    # ... which skipping it would not
        return
    G = od.compute_G(z[mode + "_H"], z[mode + "_S"], S1, z[mode + "_Sigma_tau"], mu, T.T_ft, T.T_tf, wn)
    assert np.abs(G - z[mode + "_G_tau"]).max() > 1e-6


def test_search_failure_is_raised():
    """A flat N(mu) that never reaches the target within the bracket raises (find_mu :148-150)."""
    with pytest.raises(od.ChemicalPotentialSearchFailure):
        od.find_mu(0.0, 2.0, 1e-9, lambda mu: 1.0 if mu < 0.3 else 3.0)


def _energies():
    return np.load(os.path.join(GOLDEN, "orth_energies.npz"), allow_pickle=False)


@pytest.mark.parametrize("mode", MODES)
def test_compute_energy_matches_stored_energies(mode):
    """oracle.gw_oracle.compute_energy on the stored first iteration reproduces the energies the
    reference wrote for it (iter1/Energy_1b; iter1/Energy_HF minus HF/Energy_nuc; Energy_2b = 0 with
    Sigma(tau) = 0).  All four bases give the same energies (the reference's own check, 1e-10)."""
    from oracle import gw_oracle as go

    z, T, _ = _fixture()
    e = _energies()
    G, S1, St, H = (z[mode + k] for k in ("_G_tau", "_Sigma1", "_Sigma_tau", "_H"))
    e1, ehf, ec = go.compute_energy(G, S1, St, H, T.T_ft, T.T_tf, nk=1, ns=2)
    assert abs(e1 - float(e[mode + "_Energy_1b"])) < 1e-12
    assert abs(ehf + float(e[mode + "_E_nuc"]) - float(e[mode + "_Energy_HF"])) < 1e-12
    assert abs(ec - float(e[mode + "_Energy_2b"])) < 1e-14


def test_compute_energy_planted_mistakes():
    """The density matrix with the wrong sign (dm = +G(beta)) or half the one-body H in E_HF miss the
    stored values."""
    from oracle import gw_oracle as go

    z, T, _ = _fixture()
    e = _energies()
    G, S1, St, H = (z["none" + k] for k in ("_G_tau", "_Sigma1", "_Sigma_tau", "_H"))
    e1, ehf, _ = go.compute_energy(-G, S1, St, H, T.T_ft, T.T_tf, nk=1, ns=2)
    assert abs(e1 - float(e["none_Energy_1b"])) > 1.0
    e1, ehf, _ = go.compute_energy(G, S1, St, 0.5 * H, T.T_ft, T.T_tf, nk=1, ns=2)
    assert abs(ehf + float(e["none_E_nuc"]) - float(e["none_Energy_HF"])) > 0.1


def test_compute_energy_corr_equals_gm_energy():
    """The correlation part of compute_energy is gm_energy (closed form pinned in test_oracle.py)."""
    from oracle import gw_oracle as go

    rng = np.random.default_rng(3)
    z, T, _ = _fixture()
    G = z["none_G_tau"]
    St = 0.01 * (rng.standard_normal(G.shape) + 1j * rng.standard_normal(G.shape))
    _, _, ec = go.compute_energy(G, z["none_Sigma1"], St, z["none_H"], T.T_ft, T.T_tf, nk=1, ns=2)
    assert abs(ec - go.gm_energy(G, St, T.T_ft, T.T_tf, 1, 2)) < 1e-15
