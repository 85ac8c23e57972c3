"""Unit parity of the HIP building blocks (needs an MI355X): the FP64 MFMA GEMM and the
intermediates P0(tau, q) and P-tilde(tau, q) against the oracle."""
import numpy as np
import pytest

from gwhip import lib as gl
from helpers import CONFIG0, rel_err, synthetic
from oracle import gw_oracle as go

pytestmark = pytest.mark.gpu


def test_zgemm_identity_probe():
    """A = I makes C = B; a wrong MFMA lane/register mapping shows up as a permutation."""
    n = 16
    A = np.eye(n, dtype=complex)
    B = (np.arange(n)[:, None] * 100 + np.arange(n)[None, :]).astype(complex)
    C = gl.debug_zgemm(A, B)
    if not np.array_equal(C, B):
        bad = np.argwhere(C != B)
        raise AssertionError(f"layout mismatch at {len(bad)} entries; first: C{tuple(bad[0])}="
                             f"{C[tuple(bad[0])]} expected {B[tuple(bad[0])]}")


def _lanes_of(x):
    """The powers of two in a sum of at most four equal-or-distinct powers of two, as exponents."""
    out = []
    for _ in range(4):
        if x <= 0:
            break
        e = int(np.floor(np.log2(x)))
        out.append(e)
        x -= 2.0 ** e
    assert x == 0.0
    return out


def test_mfma_f64_4x4x4_lane_layout():
    """v_mfma_f64_4x4x4_4b, the instruction of the Sigma epilogue for nao <= 32: D of lane L must be
    sum_k A[16 k + 4 b + r] B[16 k + 4 b + c] with L = 16 r + 4 b + c (A_b[r][k] at lane 16 k + 4 b + r,
    B_b[k][c] at lane 16 k + 4 b + c, D_b[r][c] at lane 16 r + 4 b + c).  Probe inputs
    A = 2^(l % 16), B = 2^(16 (l % 16)) make each D a sum of 2^(a + 16 b) over its four (A lane, B lane)
    pairs; A = l / 16 + 1 (or B) with the other 1 makes D = 1 + 2 + 3 + 4 when k runs over l / 16."""
    o = gl.debug_mfma4_layout()
    assert np.all(o[1] == 10.0) and np.all(o[2] == 10.0)
    for L in range(64):
        r, blk, c = L // 16, (L // 4) % 4, L % 4
        a_lanes = [(16 * k + 4 * blk + r) % 16 for k in range(4)]
        b_lanes = [(16 * k + 4 * blk + c) % 16 for k in range(4)]
        expect = sum(2.0 ** (a + 16 * b) for a, b in zip(a_lanes, b_lanes))
        assert o[0][L] == expect, (L, _lanes_of(o[0][L]))


@pytest.mark.parametrize("shape", [(1, 1, 1), (16, 16, 4), (64, 64, 16), (70, 130, 37), (300, 257, 41)])
@pytest.mark.parametrize("conj", [(False, False), (True, False), (False, True)])
@pytest.mark.parametrize("tile", [0, 1, 2])
def test_zgemm_random(shape, conj, tile):
    M, N, K = shape
    rng = np.random.default_rng(M * 1000 + N + K)
    A = rng.standard_normal((M, K)) + 1j * rng.standard_normal((M, K))
    B = rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N))
    C = gl.debug_zgemm(A, B, *conj, shape=tile)
    Aop = np.conj(A) if conj[0] else A
    Bop = np.conj(B) if conj[1] else B
    ref = Aop @ Bop
    assert np.abs(C - ref).max() < 1e-13 * max(1.0, np.abs(ref).max()) * np.sqrt(K)


def _setup(G, V, T, d):
    k = gl.GwHipKernel(d["nk"], d["nao"], d["nso"], d["ns"], d["NQ"], T.nts, T.nw_b, d["beta"], mesh=d["mesh"])
    k.set_grid(T.T_bt, T.T_tb)
    k.upload_V(V)
    return k


def _oracle_P0(G, V, d):
    nts = G.shape[0]
    mesh = go.mesh_shape(d["nk"], d["mesh"])
    out = []
    for q in range(d["nk"]):
        P0 = np.zeros((nts, d["NQ"], d["NQ"]), complex)
        go.eval_P0_tilde(P0, G, V, q, d["nk"], d["nao"], d["nso"], d["ns"], d["NQ"], d["nso"] != d["nao"], mesh)
        go.symmetrize_P0(P0, d["nk"])
        out.append(P0)
    return np.stack(out)


@pytest.mark.parametrize("case", ["config0", "ragged"])
def test_stage_P0_and_P_tilde(case):
    if case == "config0":
        c = CONFIG0
        G, V, T, d = synthetic(c["nk"], c["nao"], c["NQ"], beta=c["beta"], lam=c["lam"], mesh=c["mesh"])
    else:
        G, V, T, d = synthetic(3, 3, 5, ns=2, beta=20.0, lam=1e3, mesh=(3, 1, 1), seed=4, sigma=0.3)
    half = T.nts // 2
    P0_ref = _oracle_P0(G, V, d)
    k = _setup(G, V, T, d)
    try:
        k.run_stages(G, 1)
        P0 = k.P_tilde()
        k.run_stages(G, 2)
        Pt = k.P_tilde()
    finally:
        k.close()
    assert rel_err(P0[:, :half], P0_ref[:, :half]) < 1e-12
    Pt_ref = np.stack([go.eval_P_tilde(P0_ref[q], T.T_bt, T.T_tb) for q in range(d["nk"])])
    assert rel_err(Pt, Pt_ref) < 1e-10
