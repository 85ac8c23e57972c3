"""The headline workload at full size: 4x4x4 k mesh (nk = 64), nao = 26, naux = 300, ns = 1,
beta = 100, IR Lambda = 1e4 (nts = 100), seed 0 -- needs an MI355X.

FP64 (the plain form):
* per q, for q = 0 and q = 37: P0(tau, q) after symmetrize_P0, P-tilde(tau, q) and q's contribution to
  Sigma against the oracle's eval_P0_tilde / symmetrize_P0 / eval_P_tilde / eval_selfenergy
  (src/gw_cpu_kernel.cpp:74-373) on the same seeded inputs, within 1e-10 relative (max-abs difference
  over max |reference|).  The oracle reads V pair by pair from the same counters the device
  generator uses (their values agree to the last bits of log / cos);
* the whole evaluation is bitwise repeatable and equals the sum of its per-q contributions.
Split precision at its default 4 slices (the precision variant of the same workload): Sigma within
1e-8 relative of FP64, and the Galitskii-Migdal energy of both (device compute_energy) within 1e-8
relative.
"""
import time

import numpy as np
import pytest

from gwhip import inputs, irgrid
from gwhip import lib as gl
from helpers import rel_err
from oracle import gw_oracle as go

pytestmark = pytest.mark.gpu

W = dict(nk=64, mesh=(4, 4, 4), nao=26, NQ=300, ns=1, beta=100.0, lam=1e4, seed=0, sigma=0.1)
TOL_FP64 = 1e-10


@pytest.fixture(scope="module")
def system():
    T = irgrid.Transformer.from_grid(irgrid.load_grid(W["lam"]), W["beta"])
    G = inputs.generate_G(W["seed"], W["nk"], W["nao"], W["ns"], T.tau, W["beta"])
    k = gl.GwHipKernel(W["nk"], W["nao"], W["nao"], W["ns"], W["NQ"], T.nts, T.nw_b, W["beta"], mesh=W["mesh"])
    k.set_grid(T.T_bt, T.T_tb)
    k.generate_V(W["seed"], W["sigma"])
    V = inputs.LazyPackedV(W["seed"], W["nk"], W["NQ"], W["nao"], W["sigma"])
    print(f"config 1: nts {T.nts} nw_b {T.nw_b}", flush=True)
    yield G, V, T, k
    k.close()


@pytest.mark.parametrize("q", [0, 37])
def test_fp64_stages_of_one_q(system, q):
    G, V, T, k = system
    nk, nao, NQ, mesh = W["nk"], W["nao"], W["NQ"], W["mesh"]
    t0 = time.time()
    r = k.eval_q(G, q, gl.PREC_FP64)
    t1 = time.time()
    P0 = np.zeros((T.nts, NQ, NQ), complex)
    go.eval_P0_tilde(P0, G, V, q, nk, nao, nao, 1, NQ, False, mesh)
    go.symmetrize_P0(P0, nk)
    e_p0 = rel_err(r["P0"], P0)
    print(f"q {q}: device {t1 - t0:.1f} s; P0 {e_p0:.2e} (oracle {time.time() - t1:.1f} s)", flush=True)
    Pt = go.eval_P_tilde(P0.copy(), T.T_bt, T.T_tb)
    e_pt = rel_err(r["P_tilde"], Pt)
    print(f"q {q}: P-tilde {e_pt:.2e}", flush=True)
    S = np.zeros_like(G)
    go.eval_selfenergy(S, G, V, Pt, q, nk, nao, nao, 1, NQ, False, mesh)
    S /= nk
    e_s = rel_err(r["Sigma_q"], S)
    print(f"q {q}: Sigma contribution {e_s:.2e} (oracle total {time.time() - t1:.1f} s)", flush=True)
    assert e_p0 < TOL_FP64 and e_pt < TOL_FP64 and e_s < TOL_FP64


def test_fp64_repeatable_and_sum_of_q(system):
    import torch

    G, V, T, k = system
    dev = torch.device("cuda:0")
    G_d = torch.from_numpy(G.view(np.float64)).to(dev)
    S1 = torch.empty_like(G_d)
    S2 = torch.empty_like(G_d)
    torch.cuda.synchronize()
    k.solve_device(G_d.data_ptr(), S1.data_ptr(), gl.PREC_FP64)
    k.solve_device(G_d.data_ptr(), S2.data_ptr(), gl.PREC_FP64)
    a = S1.cpu().numpy().view(np.complex128)
    assert np.array_equal(a, S2.cpu().numpy().view(np.complex128))
    part = sum(k.eval_q(G, q, gl.PREC_FP64, want=("Sigma_q",))["Sigma_q"] for q in (0, 1, 2, 3))
    rest = sum(k.eval_q(G, q, gl.PREC_FP64, want=("Sigma_q",))["Sigma_q"] for q in range(4, W["nk"]))
    print(f"whole evaluation vs sum of per-q contributions: {rel_err(part + rest, a):.2e}", flush=True)
    assert rel_err(part + rest, a) < 1e-13


def test_split_vs_fp64_and_gm_energy(system):
    import torch

    G, V, T, k = system
    dev = torch.device("cuda:0")
    G_d = torch.from_numpy(G.view(np.float64)).to(dev)
    S64 = torch.empty_like(G_d)
    Ssp = torch.empty_like(G_d)
    torch.cuda.synchronize()
    k.solve_device(G_d.data_ptr(), S64.data_ptr(), gl.PREC_FP64)
    k.set_split_slices(4)  # the default
    k.solve_device(G_d.data_ptr(), Ssp.data_ptr(), gl.PREC_SPLIT)
    a = S64.cpu().numpy().view(np.complex128)
    b = Ssp.cpu().numpy().view(np.complex128)
    err = rel_err(b, a)
    wn = (2 * T.n_fermi + 1) * np.pi / W["beta"]
    d = gl.Dyson(T.nts, len(wn), 1, W["nk"], W["nk"], W["nao"], W["nao"])
    try:
        d.set_grid(T.T_ft, T.T_tf, wn)
        z = np.zeros((1, W["nk"], W["nao"], W["nao"]), complex)
        d.set_system(z, z + np.eye(W["nao"]))
        e64 = d.energy_device(G_d.data_ptr(), torch.zeros(2 * z.size, dtype=torch.float64, device=dev).data_ptr(),
                              S64.data_ptr())[2]
        esp = d.energy_device(G_d.data_ptr(), torch.zeros(2 * z.size, dtype=torch.float64, device=dev).data_ptr(),
                              Ssp.data_ptr())[2]
    finally:
        d.close()
    e_gm = abs(esp - e64) / abs(e64)
    print(f"split vs FP64: Sigma {err:.2e}, E_corr {e64!r} vs {esp!r} ({e_gm:.2e})", flush=True)
    assert 0.0 < err < 1e-8
    assert e_gm < 1e-8
