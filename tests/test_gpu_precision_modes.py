"""Per-contraction precision (the reference's P_sp / Sigma_sp, src/green/mbpt/common_defs.h:162-163)
and the leakage inputs (P0 and P-tilde of the first q, src/gw_cpu_kernel.cpp:237-239,276-278) --
needs an MI355X.

* GWHIP_PREC_SPLIT_SIGMA leaves the P0 stage in FP64, so its P-tilde is bit-identical to the FP64
  evaluation's; GWHIP_PREC_SPLIT_P0 leaves the Sigma stage in FP64.  Each mode differs from FP64 (the
  split products are in use) and stays within 1e-8 of the oracle.
* gwhip_get_leakage_inputs returns the oracle's P0(tau, q = 0) (after symmetrize_P0, full tau grid)
  and P-tilde(tau, q = 0) to 1e-10.
"""
import numpy as np
import pytest

from gwhip import lib as gl
from helpers import CONFIG0, rel_err, synthetic
from oracle import gw_oracle as go

pytestmark = pytest.mark.gpu


def _setup():
    c = CONFIG0
    G, V, T, d = synthetic(c["nk"], c["nao"], c["NQ"], ns=c["ns"], beta=c["beta"], lam=c["lam"], mesh=c["mesh"])
    k = gl.GwHipKernel(d["nk"], d["nao"], d["nso"], d["ns"], d["NQ"], T.nts, T.nw_b, d["beta"], mesh=d["mesh"])
    k.set_grid(T.T_bt, T.T_tb)
    k.upload_V(V)
    return G, V, T, d, k


def test_per_contraction_precision():
    G, V, T, d, k = _setup()
    try:
        out = {}
        for p in (gl.PREC_FP64, gl.PREC_SPLIT, gl.PREC_SPLIT_P0, gl.PREC_SPLIT_SIGMA):
            S = k.solve(G, p)
            out[p] = (S, k.P_tilde())
        with pytest.raises(gl.GwhipError) as e:
            k.solve(G, 4)
        assert e.value.status == gl.ERR_ARG
    finally:
        k.close()
    S_ref = go.gw_sigma(G, V, T.T_bt, T.T_tb, d["nk"], d["nao"], d["nso"], d["ns"], d["NQ"], mesh=d["mesh"])
    S64, P64 = out[gl.PREC_FP64]
    assert rel_err(S64, S_ref) < 1e-10
    assert np.array_equal(out[gl.PREC_SPLIT_SIGMA][1], P64)          # P0 stage in FP64
    assert not np.array_equal(out[gl.PREC_SPLIT_P0][1], P64)         # P0 stage split
    for p in (gl.PREC_SPLIT, gl.PREC_SPLIT_P0, gl.PREC_SPLIT_SIGMA):
        err = rel_err(out[p][0], S_ref)
        print(f"precision {p}: Sigma {err:.2e} vs oracle", flush=True)
        assert 0.0 < err < 1e-8
    assert not np.array_equal(out[gl.PREC_SPLIT_P0][0], out[gl.PREC_SPLIT_SIGMA][0])


@pytest.mark.parametrize("precision", [gl.PREC_FP64, gl.PREC_SPLIT])
def test_leakage_inputs(precision):
    G, V, T, d, k = _setup()
    try:
        with pytest.raises(gl.GwhipError) as e:
            k.leakage_inputs()
        assert e.value.status == gl.ERR_STATE
        k.set_leakage_capture(True)
        k.solve(G, precision)
        P0, Pt = k.leakage_inputs()
    finally:
        k.close()
    NQ = d["NQ"]
    P0_ref = np.zeros((T.nts, NQ, NQ), complex)
    go.eval_P0_tilde(P0_ref, G, V, 0, d["nk"], d["nao"], d["nso"], d["ns"], NQ, False, d["mesh"])
    go.symmetrize_P0(P0_ref, d["nk"])
    Pt_ref = go.eval_P_tilde(P0_ref.copy(), T.T_bt, T.T_tb)
    tol = 1e-10 if precision == gl.PREC_FP64 else 1e-8
    print(f"leakage inputs: P0 {rel_err(P0, P0_ref):.2e}, P-tilde {rel_err(Pt, Pt_ref):.2e}", flush=True)
    assert rel_err(P0, P0_ref) < tol
    assert rel_err(Pt, Pt_ref) < tol
