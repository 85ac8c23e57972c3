"""Galitskii-Migdal, one-body and Hartree-Fock energies on the device (gwhip_dyson_energy) -- needs an
MI355X.

compute_energy (src/common_utils.cpp:5-55, called by dyson::diff, src/dyson.cpp:290-309) through the
C ABI:
* the reference's four HF_orth_invariance first iterations: E_1b and E_HF (less the stored nuclear
  repulsion) equal the values the reference wrote (iter1/Energy_1b, iter1/Energy_HF) to 1e-12, and
  E_corr = 0 (their Sigma(tau) is zero);
* random systems with a non-zero Sigma(tau) -- ns = 1, ns = 2 with weighted irreducible points, X2C --
  against oracle.gw_oracle.compute_energy (pinned by those records and the closed form) to 1e-12
  relative;
* the GW self-energy of the nk = 8, nao = 8, naux = 40 system, split precision against FP64: the
  correlation-energy error the split products leave.
"""
import os

import numpy as np
import pytest

from gwhip import lib as gl
from helpers import CONFIG0, GOLDEN, rel_err, synthetic, transformer
from oracle import gw_oracle as go

pytestmark = pytest.mark.gpu

MODES = ("none", "lowdin", "mo", "natural")


def _orth():
    z = np.load(os.path.join(GOLDEN, "dyson_orth.npz"), allow_pickle=False)
    e = np.load(os.path.join(GOLDEN, "orth_energies.npz"), allow_pickle=False)
    T = transformer(float(z["lam"]), float(z["beta"]), z["tau"])
    wn = (2 * T.n_fermi + 1) * np.pi / float(z["beta"])
    return z, e, T, wn


def _dyson(T, wn, ns, nk, ink, nao, nso):
    d = gl.Dyson(T.nts, len(wn), ns, nk, ink, nao, nso)
    d.set_grid(T.T_ft, T.T_tf, wn)
    return d


@pytest.mark.parametrize("mode", MODES)
def test_orth_records(mode):
    z, e, T, wn = _orth()
    d = _dyson(T, wn, 2, 1, 1, 10, 10)
    try:
        d.set_system(z[mode + "_H"], z[mode + "_S"])
        e1, ehf, ec = d.energy(z[mode + "_G_tau"], z[mode + "_Sigma1"], z[mode + "_Sigma_tau"])
    finally:
        d.close()
    print(f"{mode}: E_1b {e1!r} E_HF {ehf!r} E_corr {ec!r}", flush=True)
    assert abs(e1 - float(e[mode + "_Energy_1b"])) < 1e-12
    assert abs(ehf + float(e[mode + "_E_nuc"]) - float(e[mode + "_Energy_HF"])) < 1e-12
    assert abs(ec) < 1e-15


@pytest.mark.parametrize("case", [dict(ns=1, nk=4, ink=4, nao=7, nso=7, w=None),
                                  dict(ns=2, nk=8, ink=3, nao=6, nso=6, w=[1.0, 3.0, 4.0]),
                                  dict(ns=1, nk=2, ink=2, nao=5, nso=10, w=None)])
def test_random_systems(case):
    z, e, T, wn = _orth()
    rng = np.random.default_rng(11)
    c = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    h = lambda x: 0.5 * (x + np.conj(np.swapaxes(x, -1, -2)))
    ns, ink, nso = case["ns"], case["ink"], case["nso"]
    H = h(c(ns, ink, nso, nso))
    S = np.eye(nso) + 0.1 * h(c(ns, ink, nso, nso)) / np.sqrt(nso)
    G = h(c(T.nts, ns, ink, nso, nso))
    S1 = h(c(ns, ink, nso, nso))
    St = 0.05 * h(c(T.nts, ns, ink, nso, nso))
    d = _dyson(T, wn, ns, case["nk"], ink, case["nao"], nso)
    try:
        d.set_system(H, S, case["w"])
        got = d.energy(G, S1, St)
    finally:
        d.close()
    ref = go.compute_energy(G, S1, St, H, T.T_ft, T.T_tf, case["nk"], ns, X2C=nso != case["nao"], weights=case["w"])
    print(f"{case}: {got} vs {ref}", flush=True)
    for a, b in zip(got, ref):
        assert abs(a - b) < 1e-12 * max(1.0, abs(b))


def test_split_correlation_energy_config0():
    """E_corr of the FP64 and split-precision Sigma at the CPU-runnable workload (nk 8, nao 8, naux 40):
    the device energy of the FP64 Sigma equals the oracle's gm_energy, and the split Sigma's energy is
    within 1e-8 relative of it."""
    c = CONFIG0
    G, V, T, d = synthetic(c["nk"], c["nao"], c["NQ"], ns=c["ns"], beta=c["beta"], lam=c["lam"], mesh=c["mesh"])
    k = gl.GwHipKernel(d["nk"], d["nao"], d["nso"], d["ns"], d["NQ"], T.nts, T.nw_b, d["beta"], mesh=d["mesh"])
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.upload_V(V)
        S64 = k.solve(G, gl.PREC_FP64)
        Ssp = k.solve(G, gl.PREC_SPLIT)
    finally:
        k.close()
    wn = (2 * T.n_fermi + 1) * np.pi / d["beta"]
    dy = _dyson(T, wn, d["ns"], d["nk"], d["nk"], d["nao"], d["nso"])
    S1 = np.zeros(G.shape[1:], complex)
    H = np.zeros_like(S1)
    try:
        dy.set_system(H, np.broadcast_to(np.eye(d["nso"]), S1.shape).copy())
        e64 = dy.energy(G, S1, S64)[2]
        esp = dy.energy(G, S1, Ssp)[2]
    finally:
        dy.close()
    ref = go.gm_energy(G, S64, T.T_ft, T.T_tf, d["nk"], d["ns"])
    print(f"E_corr fp64 {e64!r} oracle {ref!r} split {esp!r}; rel {abs(esp - e64) / abs(e64):.2e}", flush=True)
    assert abs(e64 - ref) < 1e-12 * abs(ref)
    assert abs(esp - e64) < 1e-8 * abs(e64)
    assert rel_err(Ssp, S64) < 1e-8
