"""The Dyson step's definiteness guard -- needs an MI355X.

For the polarisation of a physical G, I - P0(i Omega) is Hermitian positive definite, and the device
inverts it without pivoting (the reference factorises it with a pivoted LDLT,
src/gw_cpu_kernel.cpp:245-250).  A G that is not physical can make I - P0 indefinite; the device
then reports GWHIP_ERR_NOT_POSITIVE_DEFINITE instead of returning a degraded Sigma.

The indefinite case is made on the host: G' = i G turns P0 (bilinear in G) into -P0, and with
V ~ N(0, 1) the smallest eigenvalue of I - P0'(i Omega) is about -100, which the test confirms with
the oracle on the CPU before the device sees it.  Every kernel exits normally either way.
"""
import numpy as np
import pytest

from gwhip import lib as gl
from helpers import rel_err, synthetic
from oracle import gw_oracle as go

pytestmark = pytest.mark.gpu

NK, MESH, NAO, NQ = 2, (2, 1, 1), 4, 12


def _system():
    return synthetic(NK, NAO, NQ, ns=1, beta=20.0, lam=1e3, mesh=MESH, seed=5, sigma=1.0)


def _min_eig_I_minus_P0(G, V, T):
    lo = np.inf
    for q in range(NK):
        P0 = np.zeros((T.nts, NQ, NQ), complex)
        go.eval_P0_tilde(P0, G, V, q, NK, NAO, NAO, 1, NQ, False, MESH)
        go.symmetrize_P0(P0, NK)
        for p in np.tensordot(T.T_bt, P0, axes=(1, 0)):
            A = np.eye(NQ) - p
            lo = min(lo, np.linalg.eigvalsh(0.5 * (A + np.conj(A.T))).min())
    return lo


def _kernel(T, d):
    k = gl.GwHipKernel(d["nk"], d["nao"], d["nso"], d["ns"], d["NQ"], T.nts, T.nw_b, d["beta"], mesh=d["mesh"])
    return k


def test_physical_G_passes_the_guard():
    G, V, T, d = _system()
    assert _min_eig_I_minus_P0(G, V, T) > 0.1
    k = _kernel(T, d)
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.upload_V(V)
        S = k.solve(G)
    finally:
        k.close()
    S_ref = go.gw_sigma(G, V, T.T_bt, T.T_tb, NK, NAO, NAO, 1, NQ, mesh=MESH)
    assert rel_err(S, S_ref) < 1e-10


@pytest.mark.parametrize("precision", [gl.PREC_FP64, gl.PREC_SPLIT])
def test_indefinite_I_minus_P0_is_reported(precision):
    G, V, T, d = _system()
    Gx = 1j * G
    assert _min_eig_I_minus_P0(Gx, V, T) < -10.0      # indefinite, confirmed on the CPU
    k = _kernel(T, d)
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.upload_V(V)
        with pytest.raises(gl.GwhipError) as e:
            k.solve(Gx, precision)
        assert e.value.status == gl.ERR_NOT_POSITIVE_DEFINITE
        with pytest.raises(gl.GwhipError) as e:
            k.vjp(Gx, np.ones_like(Gx))
        assert e.value.status == gl.ERR_NOT_POSITIVE_DEFINITE
        S = k.solve(G, precision)                       # the flag does not stick to the next call
    finally:
        k.close()
    S_ref = go.gw_sigma(G, V, T.T_bt, T.T_tb, NK, NAO, NAO, 1, NQ, mesh=MESH)
    assert rel_err(S, S_ref) < (1e-10 if precision == gl.PREC_FP64 else 1e-8)
