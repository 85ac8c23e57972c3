"""Device vector-Jacobian product of Sigma[G] -- needs an MI355X.

Checks, all through the C ABI (and once through the torch.autograd harness):
* against torch autograd of the dense FP64 einsum forward (oracle/gw_vjp.py) on small systems
  (ns = 1, ns = 2, X2C) and on the nk = 8, nao = 8, naux = 40 system: 1e-10 relative;
* stored (P0/P-tilde kept by the forward) and recomputed modes agree;
* the GEMM form of the orbital products (tuning key orbital_gemm; always so for nao > 48) on the
  small systems and at nao = 50: the same 1e-10 against autograd;
* at nk = 27 (3x3x3), nao = 26, naux = 250, beta = 100, IR 1e4 with a Hermitian-per-tau N(0,1)
  cotangent: central finite differences (step 1e-6) of the device FP64 forward along random
  directions agree with Re<G_bar, dG>, and both modes are timed.
"""
import time

import numpy as np
import pytest

from gwhip import inputs, irgrid
from gwhip import lib as gl
from helpers import CONFIG0, rel_err, synthetic
from oracle import gw_vjp as gv

pytestmark = pytest.mark.gpu

SMALL = [
    dict(nk=2, mesh=(2, 1, 1), nao=3, NQ=5, ns=1),
    dict(nk=3, mesh=(3, 1, 1), nao=2, NQ=4, ns=2),
    dict(nk=2, mesh=(2, 1, 1), nao=2, NQ=4, ns=1, nso=4),
    dict(nk=4, mesh=(2, 2, 1), nao=17, NQ=33, ns=1),
]


def _herm_noise(shape, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return 0.5 * (x + np.conj(np.swapaxes(x, -1, -2)))


def _kernel(d, T):
    k = gl.GwHipKernel(d["nk"], d["nao"], d["nso"], d["ns"], d["NQ"], T.nts, T.nw_b, d["beta"], mesh=d["mesh"])
    k.set_grid(T.T_bt, T.T_tb)
    return k


def _compare(G, V, T, d, seed=11, gemm=False):
    Sbar = _herm_noise(G.shape, seed)
    args = (V, T.T_bt, T.T_tb, d["nk"], d["nao"], d["nso"], d["ns"], d["NQ"], d["mesh"])
    _, g_ref = gv.vjp_autograd(G, Sbar, *args)
    k = _kernel(d, T)
    try:
        k.upload_V(V)
        if gemm:
            k.set_tuning(orbital_gemm=1)
        g_rec = k.vjp(G, Sbar, gl.VJP_RECOMPUTE)
        k.set_vjp_store(True)
        k.solve(G)
        g_sto = k.vjp(G, Sbar, gl.VJP_STORED)
    finally:
        k.close()
    return g_ref, g_rec, g_sto


@pytest.mark.parametrize("ci", range(len(SMALL)))
def test_vjp_small_cases(ci):
    c = SMALL[ci]
    nso = c.get("nso", c["nao"])
    G, V, T, d = synthetic(c["nk"], c["nao"], c["NQ"], ns=c["ns"], beta=20.0, lam=1e3, mesh=c["mesh"], seed=5,
                           nso=nso, sigma=0.3)
    g_ref, g_rec, g_sto = _compare(G, V, T, d)
    print(f"case {c}: recompute {rel_err(g_rec, g_ref):.2e} stored {rel_err(g_sto, g_ref):.2e}")
    assert rel_err(g_rec, g_ref) < 1e-10
    assert rel_err(g_sto, g_ref) < 1e-10


@pytest.mark.parametrize("c", SMALL[:3] + [dict(nk=2, mesh=(2, 1, 1), nao=50, NQ=6, ns=1)])
def test_vjp_gemm_form(c):
    nso = c.get("nso", c["nao"])
    G, V, T, d = synthetic(c["nk"], c["nao"], c["NQ"], ns=c["ns"], beta=20.0, lam=1e3, mesh=c["mesh"], seed=5,
                           nso=nso, sigma=0.3)
    g_ref, g_rec, g_sto = _compare(G, V, T, d, gemm=True)
    print(f"GEMM form {c}: recompute {rel_err(g_rec, g_ref):.2e} stored {rel_err(g_sto, g_ref):.2e}", flush=True)
    assert rel_err(g_rec, g_ref) < 1e-10
    assert rel_err(g_sto, g_ref) < 1e-10


def test_vjp_config0():
    c = CONFIG0
    G, V, T, d = synthetic(c["nk"], c["nao"], c["NQ"], beta=c["beta"], lam=c["lam"], mesh=c["mesh"])
    g_ref, g_rec, g_sto = _compare(G, V, T, d)
    print(f"nk=8 nao=8 naux=40: recompute {rel_err(g_rec, g_ref):.2e} stored {rel_err(g_sto, g_ref):.2e}")
    assert rel_err(g_rec, g_ref) < 1e-10
    assert rel_err(g_sto, g_rec) < 1e-12


def test_stored_mode_needs_a_kept_forward():
    G, V, T, d = synthetic(2, 3, 5, beta=20.0, lam=1e3, mesh=(2, 1, 1), seed=2)
    k = _kernel(d, T)
    try:
        k.upload_V(V)
        k.solve(G)  # not kept
        with pytest.raises(gl.GwhipError) as e:
            k.vjp(G, G, gl.VJP_STORED)
        assert e.value.status == gl.ERR_STATE
        with pytest.raises(gl.GwhipError) as e:
            k.vjp(G, G, 7)
        assert e.value.status == gl.ERR_ARG
    finally:
        k.close()


def test_torch_autograd_harness():
    import torch

    from gwhip import autograd as ga
    G, V, T, d = synthetic(4, 5, 12, beta=20.0, lam=1e3, mesh=(2, 2, 1), seed=8, sigma=0.3)
    Sbar = _herm_noise(G.shape, 3)
    k = _kernel(d, T)
    try:
        k.upload_V(V)
        dev = torch.device("cuda:0")
        Gt = torch.tensor(G, device=dev, requires_grad=True)
        S = ga.sigma(Gt, k)
        (g,) = torch.autograd.grad(S, Gt, grad_outputs=torch.tensor(Sbar, device=dev))
        g_direct = k.vjp(G, Sbar)
        S_direct = k.solve(G)
    finally:
        k.close()
    assert np.array_equal(S.detach().cpu().numpy(), S_direct)
    assert np.array_equal(g.cpu().numpy(), g_direct)


def test_vjp_nk27_finite_differences_and_modes():
    """nk = 27, nao = 26, naux = 250, beta = 100, IR 1e4: FD along random directions."""
    nk, mesh, nao, NQ, beta = 27, (3, 3, 3), 26, 250, 100.0
    T = irgrid.Transformer.from_grid(irgrid.load_grid(1e4), beta)
    G = inputs.generate_G(0, nk, nao, 1, T.tau, beta)
    Sbar = _herm_noise(G.shape, 1)
    d = dict(nk=nk, nao=nao, nso=nao, ns=1, NQ=NQ, beta=beta, mesh=mesh)
    k = _kernel(d, T)
    try:
        k.generate_V(0, 0.1)
        t0 = time.perf_counter()
        g_rec = k.vjp(G, Sbar, gl.VJP_RECOMPUTE)
        t_rec = time.perf_counter() - t0
        ms_rec = k.vjp_ms()
        k.set_vjp_store(True)
        t0 = time.perf_counter()
        k.solve(G)
        t_fwd = time.perf_counter() - t0
        g_sto = k.vjp(G, Sbar, gl.VJP_STORED)
        ms_sto = k.vjp_ms()
        print(f"\nnk=27 nao=26 naux=250: forward {t_fwd:.2f} s; VJP recompute {ms_rec / 1e3:.2f} s "
              f"(host {t_rec:.2f} s), stored {ms_sto / 1e3:.2f} s; stored vs recompute {rel_err(g_sto, g_rec):.2e}")
        assert rel_err(g_sto, g_rec) < 1e-12
        k.set_vjp_store(False)
        h = 1e-6
        for seed in (21, 22):
            dG = _herm_noise(G.shape, seed)
            Sp = k.solve(G + h * dG)
            Sm = k.solve(G - h * dG)
            fd = float(np.real(np.vdot(Sbar, (Sp - Sm) / (2 * h))))
            an = float(np.real(np.vdot(g_rec, dG)))
            print(f"direction {seed}: FD {fd:.12e} VJP {an:.12e} rel {abs(fd - an) / abs(an):.2e}")
            assert abs(fd - an) < 1e-6 * abs(an)
    finally:
        k.close()
