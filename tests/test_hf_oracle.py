"""Pins of the Hartree-Fock oracle (oracle/hf_oracle.py) -- CPU only.

* the reference's own HF_X2C fixture (test/solvers_test.cpp:388; nk=1, X2C, NQ=154), at the
  reference's 1e-6 and at the level reached;
* an independent einsum over the full V(k1, k2) table for ns = 1, ns = 2 and X2C on small
  random systems (with a non-zero Madelung constant);
* the scalar closed form (nao = NQ = nk = 1);
* planted mistakes, made in the oracle on the CPU, that the pins reject.
"""
import os

import numpy as np
import pytest

from helpers import GOLDEN
from oracle import gw_oracle as go
from oracle import gw_vjp as gv
from oracle import hf_oracle as ho


def _fixture():
    z = np.load(os.path.join(GOLDEN, "hf_x2c.npz"), allow_pickle=False)
    nao, nso, ns, NQ, nk = (int(z[k]) for k in ("nao", "nso", "ns", "NQ", "nk"))
    dm = ho.dm_from_G(z["G_beta"], ns, nso != nao)
    return dm, z["V"], z["S_k"], float(z["madelung"]), dict(nk=nk, nao=nao, nso=nso, ns=ns, NQ=NQ), z["Sigma1"]


def test_fixture_hf_x2c():
    dm, V, S, mad, d, F_ref = _fixture()
    F = ho.hf_fock(dm, V, S, mad, d["nk"], d["nao"], d["nso"], d["ns"], d["NQ"])
    assert np.abs(F - F_ref).max() < 1e-6     # reference tolerance
    assert np.abs(F - F_ref).max() < 1e-10    # reached: 8.5e-12


def random_system(nk, nao, nso, ns, NQ, seed):
    rng = np.random.default_rng(seed)
    npair = nk * (nk + 1) // 2
    Vp = 0.3 * (rng.standard_normal((npair, NQ, nao, nao)) + 1j * rng.standard_normal((npair, NQ, nao, nao)))
    for k in range(nk):  # V(k, k) blocks Hermitian per Q, as df_integral_t stores them
        i = go.pair_index(k, k)
        Vp[i] = 0.5 * (Vp[i] + np.conj(np.swapaxes(Vp[i], -1, -2)))
    nspin = 1 if nso != nao else ns
    x = rng.standard_normal((nspin, nk, nso, nso)) + 1j * rng.standard_normal((nspin, nk, nso, nso))
    dm = 0.5 * (x + np.conj(np.swapaxes(x, -1, -2)))
    y = rng.standard_normal((nspin, nk, nso, nso)) + 1j * rng.standard_normal((nspin, nk, nso, nso))
    S = np.eye(nso) + 0.1 * (y + np.conj(np.swapaxes(y, -1, -2)))
    return dm, Vp, S, 0.37


def brute_force(dm, Vp, S, mad, nk, nao, nso, ns, NQ):
    """Full-spinor einsum restatement: one block-diagonal V per k pair."""
    Vf = gv.v_full(Vp, nk)                                        # [k1, k2, Q, i, j]
    pns = nso // nao
    Vs = np.zeros((nk, nk, NQ, nso, nso), np.complex128)          # spin-diagonal V
    for a in range(pns):
        Vs[..., a * nao:(a + 1) * nao, a * nao:(a + 1) * nao] = Vf
    X2C = pns == 2
    pref_x = 1.0 if (ns == 2 or X2C) else 0.5
    diag = np.array([Vs[k, k] for k in range(nk)])                # [k, Q, i, j]
    J = np.einsum("kQab,skba->Q", diag, dm) / nk
    F = np.einsum("Q,kQij->kij", J, diag)[None].repeat(dm.shape[0], 0)
    F = F - pref_x * np.einsum("kpQia,spab,kpQjb->skij", Vs, dm, np.conj(Vs)) / nk
    F = F - pref_x * mad * np.einsum("skij,skjl,sklm->skim", S, dm, S)
    if X2C:  # the reference keeps the (b, a) block as the adjoint of (a, b) and S = S_aa on both sides
        Saa = S[0][:, :nao, :nao]
        out = F.copy()
        for blk in ((slice(0, nao), slice(0, nao)), (slice(nao, nso), slice(nao, nso)), (slice(0, nao), slice(nao, nso))):
            ex = -np.einsum("kpQia,pab,kpQjb->kij", Vf, dm[0][:, blk[0], blk[1]], np.conj(Vf)) / nk
            md = -mad * np.einsum("kij,kjl,klm->kim", Saa, dm[0][:, blk[0], blk[1]], Saa)
            jd = np.einsum("Q,kQij->kij", J, np.array([Vf[k, k] for k in range(nk)])) if blk[0] == blk[1] else 0
            out[0][:, blk[0], blk[1]] = jd + ex + md
        out[0][:, nao:, :nao] = np.conj(np.swapaxes(out[0][:, :nao, nao:], -1, -2))
        F = out
    return F


@pytest.mark.parametrize("case", [dict(nk=3, nao=3, nso=3, ns=1, NQ=5), dict(nk=2, nao=2, nso=2, ns=2, NQ=4),
                                  dict(nk=2, nao=2, nso=4, ns=1, NQ=3)])
def test_brute_force(case):
    dm, Vp, S, mad = random_system(case["nk"], case["nao"], case["nso"], case["ns"], case["NQ"], 3)
    args = (dm, Vp, S, mad, case["nk"], case["nao"], case["nso"], case["ns"], case["NQ"])
    F = ho.hf_fock(*args)
    assert np.abs(F - brute_force(*args)).max() < 1e-12 * max(1.0, np.abs(F).max())


def test_scalar_closed_form():
    v, d, s, mad = 0.7 - 0.2j, 0.3 + 0.0j, 1.1 + 0.0j, 0.25
    Vp = np.array([[[[v.real + 0j]]]])      # V(0,0) must be Hermitian: real scalar
    F = ho.hf_scalar(np.array([[[[d]]]]), Vp, np.array([[[[s]]]]), mad, 1, 1, 1, 1)
    expected = v.real * v.real * d - 0.5 * v.real * d * v.real - 0.5 * mad * s * d * s
    assert abs(F[0, 0, 0, 0] - expected) < 1e-15


@pytest.mark.parametrize("mistake", ["no_conj", "no_half", "madelung_sign"])
def test_planted_mistakes_are_rejected(monkeypatch, mistake):
    case = dict(nk=3, nao=3, nso=3, ns=1, NQ=5)
    dm, Vp, S, mad = random_system(case["nk"], case["nao"], case["nso"], case["ns"], case["NQ"], 5)
    args = (dm, Vp, S, mad, case["nk"], case["nao"], case["nso"], case["ns"], case["NQ"])
    ref = brute_force(*args)
    orig_einsum = np.einsum

    def patched(spec, *ops, **kw):
        if spec == "Qia,Qja->ij" and mistake == "no_conj":
            return orig_einsum(spec, ops[0], np.conj(ops[1]), **kw)
        if spec == "Qia,Qja->ij" and mistake == "no_half":
            return 2.0 * orig_einsum(spec, *ops, **kw)
        return orig_einsum(spec, *ops, **kw)

    if mistake == "madelung_sign":
        F = ho.hf_scalar(dm, Vp, S, -mad, case["nk"], case["nao"], case["ns"], case["NQ"])
    else:
        monkeypatch.setattr(ho.np, "einsum", patched)
        F = ho.hf_scalar(dm, Vp, S, mad, case["nk"], case["nao"], case["ns"], case["NQ"])
        monkeypatch.setattr(ho.np, "einsum", orig_einsum)
    assert np.abs(F - ref).max() > 1e-6
