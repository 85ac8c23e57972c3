"""Split-precision path (int8 balanced-digit slices, exact int32 accumulation, FP64 accumulation
across chunks) -- needs an MI355X.

Bars: the slice GEMM puts every product in its place (identity probe) and stays within the slicing
bound on random products; the split Sigma is within 1e-8 relative (max-abs difference over max
|Sigma|) of the FP64 oracle.
"""
import numpy as np
import pytest

from gwhip import lib as gl
from helpers import min_eig_I_minus_P0, CONFIG0, load_fixture, rel_err, synthetic
from oracle import gw_oracle as go

pytestmark = pytest.mark.gpu

TOL_SPLIT = 1e-8


@pytest.mark.parametrize("tile,n", [(2, 16), (4, 16), (4, 64), (4, 70)])
def test_split_gemm_identity_probe(tile, n):
    """A = I times distinct entries of B returns B to the slicing accuracy (radix-254 digits do not
    represent every dyadic number exactly): checks the int8 MFMA layouts of both block tiles, where a
    misplaced element would be off by O(1)."""
    A = np.eye(n, dtype=complex)
    B = (np.arange(n)[:, None] * 100 + np.arange(n)[None, :]).astype(complex) * (1 - 2j)
    C = gl.debug_zgemm_split(A, B, 4, tile)
    bad = np.argwhere(np.abs(C - B) > 1e-6 * np.abs(B).max())
    if len(bad):
        raise AssertionError(f"layout mismatch at {len(bad)} entries; first: C{tuple(bad[0])}="
                             f"{C[tuple(bad[0])]} expected {B[tuple(bad[0])]}")


@pytest.mark.parametrize("shape", [(1, 1, 1), (70, 130, 300), (300, 257, 700)])
@pytest.mark.parametrize("ns", [3, 4, 5])
@pytest.mark.parametrize("tile", [2, 4])
def test_split_gemm_random(shape, ns, tile):
    M, N, K = shape
    rng = np.random.default_rng(M + N + K + ns)
    A = rng.standard_normal((M, K)) + 1j * rng.standard_normal((M, K))
    B = rng.standard_normal((K, N)) + 1j * rng.standard_normal((K, N))
    C = gl.debug_zgemm_split(A, B, ns, tile)
    ref = A @ B
    # per-element bound: each operand keeps ns balanced radix-254 digits relative to 2^e >= its
    # row / column maximum (truncation and the dropped pairs s + t >= ns, two real products per part)
    bound = 8.0 * (ns + 1) * K * 254.0 ** (-ns) * np.abs(A).max() * np.abs(B).max()
    assert np.abs(C - ref).max() < bound


def _kernel(d, T):
    return gl.GwHipKernel(d["nk"], d["nao"], d["nso"], d["ns"], d["NQ"], T.nts, T.nw_b, d["beta"], mesh=d["mesh"])


def _split_vs_oracle(G, V, T, d, ns=5):
    k = _kernel(d, T)
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.upload_V(V)
        k.set_split_slices(ns)
        S_split = k.solve(G, gl.PREC_SPLIT)
        S_fp64 = k.solve(G, gl.PREC_FP64)
    finally:
        k.close()
    S_ref = go.gw_sigma(G, V, T.T_bt, T.T_tb, d["nk"], d["nao"], d["nso"], d["ns"], d["NQ"], mesh=d["mesh"])
    return S_split, S_fp64, S_ref


@pytest.mark.parametrize("ns", [4, 5])
def test_split_sigma_config0(ns):
    c = CONFIG0
    G, V, T, d = synthetic(c["nk"], c["nao"], c["NQ"], beta=c["beta"], lam=c["lam"], mesh=c["mesh"])
    S_split, S_fp64, S_ref = _split_vs_oracle(G, V, T, d, ns)
    assert rel_err(S_fp64, S_ref) < 1e-10
    err = rel_err(S_split, S_ref)
    print(f"split ns={ns}: rel err vs oracle {err:.3e}")
    assert err < TOL_SPLIT
    # Galitskii-Migdal energy from both
    E_ref = go.gm_energy(G, S_ref, T.T_ft, T.T_tf, d["nk"], d["ns"])
    E_split = go.gm_energy(G, S_split, T.T_ft, T.T_tf, d["nk"], d["ns"])
    assert abs(E_split - E_ref) < 1e-8 * abs(E_ref)


@pytest.mark.parametrize("case", [
    dict(nk=3, mesh=(3, 1, 1), nao=3, NQ=5, ns=2),
    dict(nk=2, mesh=(2, 1, 1), nao=3, NQ=70, ns=1, nso=6),
])
def test_split_small_cases(case):
    nso = case.get("nso", case["nao"])
    G, V, T, d = synthetic(case["nk"], case["nao"], case["NQ"], ns=case["ns"], beta=20.0, lam=1e3,
                           mesh=case["mesh"], seed=9, nso=nso, sigma=0.3)
    S_split, _, S_ref = _split_vs_oracle(G, V, T, d)
    assert rel_err(S_split, S_ref) < TOL_SPLIT


@pytest.mark.parametrize("ns", [3, 4])
def test_fewer_slices_are_looser_but_bounded(ns):
    """Three and four balanced slices keep about 24 and 32 bits per operand: looser than five but
    bounded; the measured error is printed."""
    c = CONFIG0
    G, V, T, d = synthetic(c["nk"], c["nao"], c["NQ"], beta=c["beta"], lam=c["lam"], mesh=c["mesh"])
    S_split, _, S_ref = _split_vs_oracle(G, V, T, d, ns=ns)
    err = rel_err(S_split, S_ref)
    print(f"split ns={ns}: rel err vs oracle {err:.3e}", flush=True)
    assert 1e-13 < err < (1e-7 if ns == 4 else 1e-4)


@pytest.mark.parametrize("p0_group,s_group", [(1, 16), (12, 1), (64, 64)])
def test_split_tiles_and_orders_give_identical_sigma(p0_group, s_group):
    """Any L2 tile order and the fused or separate slicing of Z produce bit-identical Sigma (exact int32
    sums, identical slices)."""
    c = CONFIG0
    G, V, T, d = synthetic(c["nk"], c["nao"], c["NQ"], beta=c["beta"], lam=c["lam"], mesh=c["mesh"])
    k = _kernel(d, T)
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.upload_V(V)
        ref = k.solve(G, gl.PREC_SPLIT)
        out = []
        for fused in (0, 1):
            k.set_tuning(p0_group=p0_group, sigma_group=s_group, p0_fused_slices=fused)
            out.append(k.solve(G, gl.PREC_SPLIT))
    finally:
        k.close()
    for o in out:
        assert np.array_equal(o, ref)


def test_p0_k_chunks_per_launch():
    """The int8 P0 GEMM over k chunks in one launch or in launches of 1, 3 or 8 chunks: the same
    exact int32 slice products, FP64 partial sums added in another order (1e-13)."""
    c = CONFIG0
    G, V, T, d = synthetic(c["nk"], c["nao"], c["NQ"], beta=c["beta"], lam=c["lam"], mesh=c["mesh"])
    k = _kernel(d, T)
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.upload_V(V)
        out = {}
        for kch in (0, 1, 3, 8):
            k.set_tuning(p0_kchunks=kch)
            out[kch] = k.solve(G, gl.PREC_SPLIT)
    finally:
        k.close()
    S_ref = go.gw_sigma(G, V, T.T_bt, T.T_tb, d["nk"], d["nao"], d["nso"], d["ns"], d["NQ"], mesh=d["mesh"])
    for kch, S in out.items():
        print(f"P0 k chunks per launch {kch}: vs one launch {rel_err(S, out[0]):.1e}, vs oracle {rel_err(S, S_ref):.1e}",
              flush=True)
        assert rel_err(S, out[0]) < 1e-13
        assert rel_err(S, S_ref) < TOL_SPLIT


def test_split_reference_fixture():
    G, V, T, d, S_fix = load_fixture("gw_x2c")
    S_split, _, S_ref = _split_vs_oracle(G, V, T, d)
    assert rel_err(S_split, S_ref) < TOL_SPLIT
    assert np.abs(S_split - S_fix).max() < 1e-6


def test_explicit_kq_maps_match_mesh_maps():
    mesh = (2, 2, 1)
    G, V, T, d = synthetic(4, 3, 6, beta=20.0, lam=1e3, mesh=mesh, seed=3, sigma=0.3)
    nk = 4
    kp = np.array([[go.k_add(k, q, mesh, +1) for k in range(nk)] for q in range(nk)], dtype=np.int32)
    km = np.array([[go.k_add(k, q, mesh, -1) for k in range(nk)] for q in range(nk)], dtype=np.int32)
    k1 = _kernel(d, T)
    k2 = _kernel(d, T)
    try:
        for k in (k1, k2):
            k.set_grid(T.T_bt, T.T_tb)
        k2.set_kq_maps(kp, km)
        for k in (k1, k2):
            k.upload_V(V)
        S1, S2 = k1.solve(G), k2.solve(G)
        with pytest.raises(gl.GwhipError) as e:
            k2.set_kq_maps(kp + 10, km)
        assert e.value.status == gl.ERR_ARG
    finally:
        k1.close()
        k2.close()
    assert np.array_equal(S1, S2)


@pytest.mark.parametrize("case", [
    dict(nk=8, mesh=(2, 2, 2), nao=8, NQ=40, herm=True),     # one row tile band, ragged
    dict(nk=4, mesh=(2, 2, 1), nao=5, NQ=130, herm=True),    # several row tiles, upper-tile skipping
    dict(nk=4, mesh=(2, 2, 1), nao=5, NQ=130, herm=False),   # every tile (G not Hermitian)
    dict(nk=2, mesh=(2, 1, 1), nao=3, NQ=70, nso=6, herm=True),  # X2C spin blocks
])
@pytest.mark.parametrize("ns", [3, 4, 5])
@pytest.mark.parametrize("tile", [2, 4])
def test_split_slice_counts_on_ragged_shapes(case, ns, tile):
    """Both int8 block tiles (32 x 64 of 16 x 16 x 64 MFMAs, 64 x 64 of 32 x 32 x 32 MFMAs) at 3, 4 and
    5 slices per operand against the oracle: the error falls by about 254 per slice (the measured
    errors are printed) and 4 slices, the default, stay within the 1e-8 bar."""
    nso = case.get("nso", case["nao"])
    G, V, T, d = synthetic(case["nk"], case["nao"], case["NQ"], beta=20.0, lam=1e3, mesh=case["mesh"], seed=4,
                           nso=nso, sigma=0.1)
    if not case["herm"]:
        rng = np.random.default_rng(6)
        G = G + 1e-3 * (rng.standard_normal(G.shape) + 1j * rng.standard_normal(G.shape))
    assert min_eig_I_minus_P0(G, V, T, d) > 0.5  # I - P0 positive definite, confirmed on the CPU
    k = _kernel(d, T)
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.upload_V(V)
        k.set_split_slices(ns)
        k.set_tuning(p0_split_tile=tile, sigma_split_tile=tile)
        S = k.solve(G, gl.PREC_SPLIT)
    finally:
        k.close()
    S_ref = go.gw_sigma(G, V, T.T_bt, T.T_tb, d["nk"], d["nao"], d["nso"], d["ns"], d["NQ"], mesh=d["mesh"])
    err = rel_err(S, S_ref)
    print(f"split ns={ns} tile={tile} {case}: rel err vs oracle {err:.3e}", flush=True)
    assert err < {3: 1e-4, 4: TOL_SPLIT, 5: 1e-10}[ns]


@pytest.mark.parametrize("ns", [4, 5])
def test_int8_block_tiles_give_identical_sigma(ns):
    """Both int8 block tiles sum the same exact int32 slice products per K chunk and convert them to
    FP64 in the same order, so Sigma is bit-identical across them (ragged NQ = 130, Hermitian G:
    upper-tile skipping in the P0 GEMM)."""
    G, V, T, d = synthetic(4, 5, 130, beta=20.0, lam=1e3, mesh=(2, 2, 1), seed=4, sigma=0.1)
    k = _kernel(d, T)
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.upload_V(V)
        k.set_split_slices(ns)
        out = []
        for tile in (2, 4):
            k.set_tuning(p0_split_tile=tile, sigma_split_tile=tile)
            out.append(k.solve(G, gl.PREC_SPLIT))
    finally:
        k.close()
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("nao", [8, 26, 40])
def test_fused_slicing_two_spin_blocks(nao):
    """Unrestricted (two collinear spin blocks): the fused P0 kernel sums Z over both blocks in
    registers before slicing, the separate pass adds the second block's Z to the first in memory;
    the sums round differently, so the slices agree up to the last digit's rounding.  Both forms
    against the oracle, and against each other within 1e-9."""
    G, V, T, d = synthetic(2, nao, 12, ns=2, beta=20.0, lam=1e3, mesh=(2, 1, 1), seed=29, sigma=0.3)
    k = _kernel(d, T)
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.upload_V(V)
        out = []
        for fused in (0, 1):
            k.set_tuning(p0_fused_slices=fused)
            out.append(k.solve(G, gl.PREC_SPLIT))
    finally:
        k.close()
    S_ref = go.gw_sigma(G, V, T.T_bt, T.T_tb, d["nk"], d["nao"], d["nso"], d["ns"], d["NQ"], mesh=d["mesh"])
    errs = [rel_err(o, S_ref) for o in out]
    print(f"nao={nao} ns=2: separate {errs[0]:.2e}, fused {errs[1]:.2e}, fused vs separate "
          f"{rel_err(out[1], out[0]):.2e}", flush=True)
    assert max(errs) < 1e-8
    assert rel_err(out[1], out[0]) < 1e-9
