"""HIP path vs the oracle, through the C ABI (needs an MI355X).

Bar: FP64 Sigma within 1e-10 relative (max-abs difference over max |Sigma|) of the oracle
on the same seeded inputs; the reference's own fixtures within its 1e-6 abs tolerance.  The
orbital-size products have two forms -- register-chained MFMA kernels (nao <= 48) and batched
FP64 GEMMs (any nao; tuning key orbital_gemm) -- and both are checked against the oracle,
including nao = 50 (> 48, GEMM form only), X2C, unrestricted and split precision; the register
form at one, two and three 16-row blocks of nao (16, 17, 33, 40, 48), and the 4 x 4 x 4 Sigma
epilogue (nao <= 32) at each of its block counts (6, 16, 17, 30, 32).
"""
import numpy as np
import pytest

from gwhip import lib as gl
from helpers import min_eig_I_minus_P0, CONFIG0, load_fixture, rel_err, synthetic
from oracle import gw_oracle as go

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-10


def _kernel(d, T):
    return gl.GwHipKernel(d["nk"], d["nao"], d["nso"], d["ns"], d["NQ"], T.nts, T.nw_b, d["beta"], mesh=d["mesh"])


def _gpu(G, V, T, d):
    k = _kernel(d, T)
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.upload_V(V)
        S = k.solve(G)
        P = k.P_tilde()
    finally:
        k.close()
    return S, P


def _oracle(G, V, T, d):
    return go.gw_sigma(G, V, T.T_bt, T.T_tb, d["nk"], d["nao"], d["nso"], d["ns"], d["NQ"], mesh=d["mesh"],
                       return_P=True)


def test_config0_parity():
    """2x2x2 mesh, nao 8, naux 40, ns 1, beta 50, IR 1e3, seed 0."""
    c = CONFIG0
    G, V, T, d = synthetic(c["nk"], c["nao"], c["NQ"], ns=c["ns"], beta=c["beta"], lam=c["lam"], mesh=c["mesh"])
    S, P = _gpu(G, V, T, d)
    S_ref, P_ref = _oracle(G, V, T, d)
    assert rel_err(P, P_ref) < TOL_FP64
    assert rel_err(S, S_ref) < TOL_FP64


@pytest.mark.parametrize("name", ["gw_x2c", "seet_dc"])
def test_reference_fixtures(name):
    G, V, T, d, S_fix = load_fixture(name)
    S, _ = _gpu(G, V, T, d)
    S_ref, _ = _oracle(G, V, T, d)
    assert rel_err(S, S_ref) < TOL_FP64
    assert np.abs(S - S_fix).max() < 1e-6


@pytest.mark.parametrize("case", [
    dict(nk=3, mesh=(3, 1, 1), nao=3, NQ=5, ns=2),             # ragged tiles everywhere, unrestricted
    dict(nk=2, mesh=(2, 1, 1), nao=3, NQ=7, ns=1, nso=6),      # X2C spin blocks, k > 1
    dict(nk=1, mesh=(1, 1, 1), nao=1, NQ=1, ns=1),             # scalar
    dict(nk=4, mesh=(2, 2, 1), nao=5, NQ=70, ns=1),            # NQ spans two GEMM tiles
    dict(nk=2, mesh=(2, 1, 1), nao=2, NQ=700, ns=1),           # Dyson panel narrowed to 8 columns
    dict(nk=1, mesh=(1, 1, 1), nao=2, NQ=1300, ns=1),          # ... and to 4
])
def test_small_cases(case):
    nso = case.get("nso", case["nao"])
    G, V, T, d = synthetic(case["nk"], case["nao"], case["NQ"], ns=case["ns"], beta=20.0, lam=1e3,
                           mesh=case["mesh"], seed=7, nso=nso, sigma=0.3)
    S, P = _gpu(G, V, T, d)
    S_ref, P_ref = _oracle(G, V, T, d)
    assert rel_err(P, P_ref) < TOL_FP64
    assert rel_err(S, S_ref) < TOL_FP64


@pytest.mark.parametrize("hermitian", [True, False])
def test_hermitian_and_general_G(hermitian):
    """Hermitian G: only the upper P0 tiles are computed and mirrored.  A G with an
    anti-Hermitian part takes the full products and the reference's (X + X^H)/2."""
    # V ~ N(0, 0.1^2): with 0.3 this seed's forced P0(beta - tau) = P0(tau) leaves I - P0 indefinite at
    # the first non-zero Omega, which the device reports instead of inverting
    G, V, T, d = synthetic(4, 5, 70, beta=20.0, lam=1e3, mesh=(2, 2, 1), seed=12, sigma=0.1)
    if not hermitian:
        rng = np.random.default_rng(5)
        G = G + 1e-3 * (rng.standard_normal(G.shape) + 1j * rng.standard_normal(G.shape))
    assert min_eig_I_minus_P0(G, V, T, d) > 0.5  # positive definite, confirmed on the CPU
    k = _kernel(d, T)
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.upload_V(V)
        S = k.solve(G)
        S_split = k.solve(G, gl.PREC_SPLIT)
    finally:
        k.close()
    S_ref, _ = _oracle(G, V, T, d)
    assert rel_err(S, S_ref) < TOL_FP64
    assert rel_err(S_split, S_ref) < 1e-8


def test_device_generator_matches_host_generator():
    G, V, T, d = synthetic(8, 4, 12, beta=50.0, lam=1e3, mesh=(2, 2, 2), seed=11)
    k = _kernel(d, T)
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.upload_V(V)
        S_up = k.solve(G)
        k.generate_V(11, 0.1)
        S_gen = k.solve(G)
    finally:
        k.close()
    assert rel_err(S_gen, S_up) < 1e-12


def test_repeatable_bitwise():
    G, V, T, d = synthetic(8, 6, 20, beta=50.0, lam=1e3, mesh=(2, 2, 2), seed=2)
    k = _kernel(d, T)
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.upload_V(V)
        S1 = k.solve(G)
        S2 = k.solve(G)
    finally:
        k.close()
    assert np.array_equal(S1, S2)


def test_state_errors():
    G, V, T, d = synthetic(1, 2, 3, beta=20.0, lam=1e3, mesh=(1, 1, 1), seed=1)
    k = _kernel(d, T)
    try:
        with pytest.raises(gl.GwhipError) as e:
            k.solve(G)
        assert e.value.status == gl.ERR_STATE
        k.set_grid(T.T_bt, T.T_tb)
        with pytest.raises(gl.GwhipError) as e:
            k.solve(G)
        assert e.value.status == gl.ERR_STATE
    finally:
        k.close()


def test_nonfinite_input_is_reported():
    G, V, T, d = synthetic(1, 2, 3, beta=20.0, lam=1e3, mesh=(1, 1, 1), seed=1)
    G[3, 0, 0, 0, 0] = np.nan
    k = _kernel(d, T)
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.upload_V(V)
        with pytest.raises(gl.GwhipError) as e:
            k.solve(G)
        assert e.value.status == gl.ERR_NONFINITE
    finally:
        k.close()


@pytest.mark.parametrize("case", [
    dict(nk=4, mesh=(2, 2, 1), nao=26, NQ=30, ns=1),              # the benchmark's nao, GEMM form forced
    dict(nk=3, mesh=(3, 1, 1), nao=7, NQ=9, ns=2),                # unrestricted, GEMM form forced
    dict(nk=2, mesh=(2, 1, 1), nao=50, NQ=20, ns=1),              # nao > 48: GEMM form only
    dict(nk=2, mesh=(2, 1, 1), nao=50, NQ=8, ns=1, nso=100),      # X2C spin blocks at nao > 48
])
def test_orbital_gemm_form(case):
    nso = case.get("nso", case["nao"])
    G, V, T, d = synthetic(case["nk"], case["nao"], case["NQ"], ns=case["ns"], beta=20.0, lam=1e3,
                           mesh=case["mesh"], seed=13, nso=nso, sigma=0.3)
    k = _kernel(d, T)
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.upload_V(V)
        if case["nao"] <= 48:
            S_mfma = k.solve(G)
            k.set_tuning(orbital_gemm=1)
        S = k.solve(G)
        P = k.P_tilde()
        S_split = k.solve(G, gl.PREC_SPLIT)
    finally:
        k.close()
    S_ref, P_ref = _oracle(G, V, T, d)
    msg = f"GEMM form {case}: P {rel_err(P, P_ref):.2e}, Sigma {rel_err(S, S_ref):.2e}, split {rel_err(S_split, S_ref):.2e}"
    if case["nao"] <= 48:
        msg += f", vs MFMA kernels {rel_err(S, S_mfma):.2e}"
        assert rel_err(S, S_mfma) < 1e-12
    print(msg, flush=True)
    assert rel_err(P, P_ref) < TOL_FP64
    assert rel_err(S, S_ref) < TOL_FP64
    assert rel_err(S_split, S_ref) < 1e-8


@pytest.mark.parametrize("case", [
    dict(nk=2, mesh=(2, 1, 1), nao=6, NQ=12, ns=2),              # 4x4x4 epilogue, two 4-blocks, unrestricted
    dict(nk=2, mesh=(2, 1, 1), nao=16, NQ=12, ns=1),             # one 16-block, no padding; four 4-blocks
    dict(nk=2, mesh=(2, 1, 1), nao=17, NQ=12, ns=1),             # two blocks, one row in the second; seven 4-blocks
    dict(nk=2, mesh=(2, 1, 1), nao=30, NQ=10, ns=1),             # eight 4-blocks, ragged
    dict(nk=2, mesh=(2, 1, 1), nao=32, NQ=8, ns=2),              # eight full 4-blocks, unrestricted
    dict(nk=2, mesh=(2, 1, 1), nao=33, NQ=10, ns=1),             # three blocks, ragged
    dict(nk=2, mesh=(2, 1, 1), nao=48, NQ=10, ns=1),             # three full blocks (largest register form)
    dict(nk=2, mesh=(2, 1, 1), nao=40, NQ=8, ns=2),              # three blocks, unrestricted
    dict(nk=1, mesh=(1, 1, 1), nao=40, NQ=6, ns=1, nso=80),      # three blocks, X2C spin blocks
])
def test_register_kernels_every_block_count(case):
    """The register-chained orbital kernels are instantiated for one, two and three 16-row blocks
    of nao, and the 4 x 4 x 4 Sigma epilogue (nao <= 32) for two, four, seven and eight 4-row
    blocks; each instantiation (Hermitian-G upper-block epilogue, fused P0 slicing for one spin
    block, the separate slicing pass otherwise) against the oracle in both precisions."""
    nso = case.get("nso", case["nao"])
    G, V, T, d = synthetic(case["nk"], case["nao"], case["NQ"], ns=case["ns"], beta=20.0, lam=1e3,
                           mesh=case["mesh"], seed=17, nso=nso, sigma=0.3)
    k = _kernel(d, T)
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.upload_V(V)
        S = k.solve(G)
        P = k.P_tilde()
        S_split = k.solve(G, gl.PREC_SPLIT)
    finally:
        k.close()
    S_ref, P_ref = _oracle(G, V, T, d)
    print(f"register kernels {case}: P {rel_err(P, P_ref):.2e}, Sigma {rel_err(S, S_ref):.2e}, "
          f"split {rel_err(S_split, S_ref):.2e}", flush=True)
    assert rel_err(P, P_ref) < TOL_FP64
    assert rel_err(S, S_ref) < TOL_FP64
    assert rel_err(S_split, S_ref) < 1e-8


@pytest.mark.parametrize("case", [
    dict(nk=2, mesh=(2, 1, 1), nao=6, NQ=12, ns=2),              # one i group, unrestricted
    dict(nk=2, mesh=(2, 1, 1), nao=17, NQ=12, ns=1),             # two i groups, one row in the second
    dict(nk=2, mesh=(2, 1, 1), nao=30, NQ=10, ns=1),             # two i groups, ragged 4-blocks
    dict(nk=2, mesh=(2, 1, 1), nao=37, NQ=8, ns=2),              # three i groups (ten 4-blocks), unrestricted
    dict(nk=2, mesh=(2, 1, 1), nao=45, NQ=8, ns=1),              # twelve 4-blocks, ragged
    dict(nk=1, mesh=(1, 1, 1), nao=40, NQ=6, ns=1, nso=80),      # X2C spin blocks
    dict(nk=2, mesh=(2, 1, 1), nao=37, NQ=48, ns=1),             # few blocks: the Q range split in three
    dict(nk=2, mesh=(2, 1, 1), nao=20, NQ=40, ns=2),             # ... in two, unrestricted
])
def test_sigma_epilogue_forms(case):
    """The three register forms of the Sigma epilogue (4x4x4 MFMAs with every 16-row group of i
    per wave, 4x4x4 with one group per wave, 16x16x4) against each other and against the oracle;
    the first form applies up to nao = 32 only and stands in for the default above it.  With few
    (k, tau) blocks and NQ >= 32 the Q loop is split over work-groups whose partial sums are added
    in order (the last two cases)."""
    nso = case.get("nso", case["nao"])
    G, V, T, d = synthetic(case["nk"], case["nao"], case["NQ"], ns=case["ns"], beta=20.0, lam=1e3,
                           mesh=case["mesh"], seed=23, nso=nso, sigma=0.3)
    k = _kernel(d, T)
    S = {}
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.upload_V(V)
        for form in (1, 2, 3):
            k.set_tuning(epi_form=form)
            S[form] = k.solve(G)
        S_split = k.solve(G, gl.PREC_SPLIT)
    finally:
        k.close()
    S_ref, _ = _oracle(G, V, T, d)
    print(f"epilogue forms {case}: " + ", ".join(f"{f}: {rel_err(S[f], S_ref):.2e}" for f in S) +
          f", 2 vs 3 {rel_err(S[2], S[3]):.2e}, split (form 3) {rel_err(S_split, S_ref):.2e}", flush=True)
    for f in S:
        assert rel_err(S[f], S_ref) < TOL_FP64
    assert rel_err(S[2], S[3]) < 1e-12 and rel_err(S[1], S[2]) < 1e-12
    assert rel_err(S_split, S_ref) < 1e-8


def test_out_of_range_tuning_is_refused():
    """Tuning values outside each key's range (epilogue form 4, int8 tile 3, aux per wave 3) are
    refused with GWHIP_ERR_ARG and leave the handle's tuning as it was."""
    G, V, T, d = synthetic(1, 2, 3, beta=20.0, lam=1e3, mesh=(1, 1, 1), seed=1)
    k = _kernel(d, T)
    try:
        before = k.tuning()
        for bad in (dict(epi_form=4), dict(sigma_split_tile=3), dict(aux_per_wave=3)):
            with pytest.raises(gl.GwhipError) as e:
                k.set_tuning(**bad)
            assert e.value.status == gl.ERR_ARG
        assert k.tuning() == before
    finally:
        k.close()
