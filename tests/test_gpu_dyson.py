"""Device Dyson step and chemical-potential search (gwhip_dyson_*) -- needs an MI355X.

Through the C ABI, against the oracle (oracle/dyson.py, pinned by the reference's HF_orth_invariance
fixtures in test_dyson_oracle.py):
* the four HF_orth_invariance fixtures: G(tau) against the stored G (reference tolerance 1e-6) and
  against the oracle at 1e-10 relative; find_mu from 0 with tolerance 1e-9 returns the stored -0.25;
* random systems with a non-zero Sigma(tau): ns = 2 with three weighted irreducible points, X2C,
  nso = 96 (the largest the LDS inverse takes) and nso = 100, 130 (eliminated in global memory):
  G and N(mu) against the oracle.
"""
import os

import numpy as np
import pytest

from gwhip import lib as gl
from helpers import GOLDEN, rel_err, transformer
from oracle import dyson as od

pytestmark = pytest.mark.gpu

MODES = ("none", "lowdin", "mo", "natural")


def _fixture():
    z = np.load(os.path.join(GOLDEN, "dyson_orth.npz"), allow_pickle=False)
    T = transformer(float(z["lam"]), float(z["beta"]), z["tau"])
    wn = (2 * T.n_fermi + 1) * np.pi / float(z["beta"])
    return z, T, wn


def _dyson(T, wn, ns, nk, ink, nao, nso):
    d = gl.Dyson(T.nts, len(wn), ns, nk, ink, nao, nso)
    d.set_grid(T.T_ft, T.T_tf, wn)
    return d


@pytest.mark.parametrize("mode", MODES)
def test_orth_fixture(mode):
    z, T, wn = _fixture()
    H, S = z[mode + "_H"], z[mode + "_S"]
    d = _dyson(T, wn, 2, 1, 1, 10, 10)
    try:
        d.set_system(H, S)
        G = d.compute_G(z[mode + "_Sigma1"], z[mode + "_Sigma_tau"], float(z[mode + "_mu"]))
        nel, mu = d.find_mu(z[mode + "_Sigma1"], z[mode + "_Sigma_tau"], 0.0, float(z[mode + "_nel"]), 1e-9)
    finally:
        d.close()
    G_or = od.compute_G(H, S, z[mode + "_Sigma1"], z[mode + "_Sigma_tau"], float(z[mode + "_mu"]), T.T_ft, T.T_tf, wn)
    err_ref = np.abs(G - z[mode + "_G_tau"]).max()
    print(f"{mode}: G {err_ref:.2e} abs vs reference, {rel_err(G, G_or):.2e} rel vs oracle; mu {mu} nel {nel}",
          flush=True)
    assert err_ref < 1e-6 and rel_err(G, G_or) < 1e-10
    if mode == "none":
        assert mu == float(z["none_mu"]) == -0.25
    assert abs(nel - 2.0) < 2e-9


def _random(ns, ink, nao, nso, seed, nts, S_scale=0.1):
    rng = np.random.default_rng(seed)
    c = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    h = lambda x: 0.5 * (x + np.conj(np.swapaxes(x, -1, -2)))
    H = h(c(ns, ink, nso, nso)) * 0.5
    S = np.eye(nso) + S_scale * h(c(ns, ink, nso, nso)) / np.sqrt(nso)
    S1 = 0.1 * c(ns, ink, nso, nso)                    # not Hermitian: the step keeps its Hermitian part
    St = 0.02 * c(nts, ns, ink, nso, nso)
    return H, S, S1, St


@pytest.mark.parametrize("case", [dict(ns=2, nk=8, ink=3, nao=6, nso=6, w=[1.0, 3.0, 4.0]),
                                  dict(ns=1, nk=4, ink=4, nao=5, nso=10, w=None),
                                  dict(ns=1, nk=1, ink=1, nao=48, nso=96, w=None),
                                  dict(ns=1, nk=2, ink=1, nao=50, nso=100, w=[2.0]),
                                  dict(ns=2, nk=1, ink=1, nao=130, nso=130, w=None)])
def test_random_systems(case):
    z, T, wn = _fixture()
    H, S, S1, St = _random(case["ns"], case["ink"], case["nao"], case["nso"], 7, T.nts)
    d = _dyson(T, wn, case["ns"], case["nk"], case["ink"], case["nao"], case["nso"])
    try:
        d.set_system(H, S, case["w"])
        G = d.compute_G(S1, St, 0.1)
        N = d.number_of_electrons(S1, St, 0.1)
    finally:
        d.close()
    G_or = od.compute_G(H, S, S1, St, 0.1, T.T_ft, T.T_tf, wn)
    e = od.eigenspectra(H, S, S1, St, T.T_ft)
    w = case["w"] if case["w"] is not None else [1.0] * case["ink"]
    N_or = od.number_of_electrons(0.1, e, wn, T.T_tf, w, case["nk"], case["ns"], X2C=case["nso"] != case["nao"])
    print(f"{case}: G {rel_err(G, G_or):.2e}, N {N:.12f} vs {N_or:.12f}", flush=True)
    assert rel_err(G, G_or) < 1e-10
    assert abs(N - N_or) < 1e-9 * max(1.0, abs(N_or))
