"""The on-disk integral reader and the irreducible-wedge maps -- CPU only.

Oracle pins (oracle/integrals.py, oracle/symmetry.py):
* the reference's nk = 27 HF fixture (test/solvers_test.cpp:387; H2, ns = 2, nao = 2, NQ = 36,
  4 irreducible k points, 196 of 378 k pairs stored): full-mesh HF from the expanded pairs and the
  unfolded density matrix, gathered at the irreducible points, against result/Sigma1;
* the reference's nk = 27 GW fixture (test/solvers_test.cpp:389) the same way through the GW
  oracle, against result/Sigma_tau at the reference's 1e-6;
* planted mistakes (in the oracle, on the CPU): conj and trans pair operations swapped, the
  transform applied as U^H X U -- each rejected by the HF fixture.
* the reference's GW_C fixture (solvers_test.cpp:390): a time-reversal wedge (14 of 27 k points,
  tr_conj = 1 on 13) with q0_treatment EXTRAPOLATE (oracle/q0.py) against result/Sigma_tau;
  planted mistakes (correction dropped, real and imaginary parts of eps swapped) rejected; the
  q -> 0 fit against an exact quadratic.
* the time-reversal branch of value_AO is pinned by GW_C only where U = I (conjugation alone);
  with tr_conj = 1 and U != I the order of U and conjugation is parity unpinned.
* the reference's own star-of-q evaluation (oracle/symmetry.py gw_sigma_star: P-tilde of the 4
  irreducible q, Sigma of the 4 irreducible k over every q of each star, P-tilde rotated by the
  auxiliary-basis matrices k_sym_transform_p0) against the GW fixture at 2.6e-11 and the GW_C
  fixture at 2.8e-11; the same evaluation without the rotation is rejected (2e-2).  The order
  "rotate, then conjugate" of eval_p0_bz_from_ibz is read from gw_cpu_kernel.cpp:281-293; no
  fixture has tr_conj = 1 with U_q != I, so that combination is parity unpinned.

Host reader (libgwhip_h5.so, no GPU needed):
* the version rule against the reference's own pass/fail strings (solvers_test.cpp:497-508) and
  its four meta.h5 files (:470-495);
* the packed integrals read from the reference's files equal the oracle's expansion bit for bit;
  the symmetry tables equal the input file's; k +- q from fractional coordinates equal the
  Monkhorst-Pack index arithmetic;
* missing files and wrong shapes give ERR_IO.
"""
import os

import numpy as np
import pytest

from gwhip import lib as gl
from helpers import GOLDEN, transformer
from oracle import gw_oracle as go
from oracle import hf_oracle as ho
from oracle import integrals as oi
from oracle import symmetry as osy

H5 = os.path.join(GOLDEN, "h5")
DF = os.path.join(H5, "gw_sym27_df")
INPUT = os.path.join(H5, "sym27_input.h5")


def _sym():
    z = np.load(os.path.join(GOLDEN, "sym27.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


def _pairs(z):
    return z["pairs_conj_pairs_list"], z["pairs_trans_pairs_list"], z["pairs_kpair_irre_list"]


def _hf_fixture(z, v_irre, transform=None):
    h = np.load(os.path.join(GOLDEN, "hf_sym27.npz"), allow_pickle=False)
    nk, nao, ns, NQ = int(z["nk"]), int(z["nao"]), int(z["ns"]), int(z["NQ"])
    Vp = oi.packed_from_irreducible(h["V"] if v_irre is None else v_irre, nk, *_pairs(z))
    dm_ibz = ho.dm_from_G(h["G_tau"][-1], ns, False)
    if transform is None:
        dm = osy.unfold(dm_ibz, z["k_bz2ibz"], z["k_ibz2bz"], z["k_k_sym_transform_ao"], z["k_tr_conj"])
    else:
        dm = np.empty((ns, nk, nao, nao), complex)
        for s in range(ns):
            for k in range(nk):
                ik = osy.reduced_point(z["k_bz2ibz"], z["k_ibz2bz"], k)
                x = dm_ibz[s, ik]
                dm[s, k] = x if z["k_ibz2bz"][ik] == k else transform(x, z["k_k_sym_transform_ao"][k])
    F = ho.hf_scalar(dm, Vp, z["S_k"], float(z["madelung"]), nk, nao, ns, NQ)
    return osy.full_to_ibz(F, z["k_ibz2bz"]), h["Sigma1"]


def test_hf_sym27_fixture():
    F, ref = _hf_fixture(_sym(), None)
    err = np.abs(F - ref).max()
    assert err < 1e-6      # reference tolerance (IsCloseTo default)
    assert err < 1e-11     # reached: 2.8e-13


def test_gw_sym27_fixture():
    z = _sym()
    g = np.load(os.path.join(GOLDEN, "gw_sym27.npz"), allow_pickle=False)
    nk, nao, ns, NQ = int(z["nk"]), int(z["nao"]), int(z["ns"]), int(z["NQ"])
    Vp = oi.packed_from_irreducible(g["V"], nk, *_pairs(z))
    G = osy.unfold(g["G_tau"], z["k_bz2ibz"], z["k_ibz2bz"], z["k_k_sym_transform_ao"], z["k_tr_conj"])
    T = transformer(float(g["lam"]), float(g["beta"]), g["tau"])
    S = go.gw_sigma(G, Vp, T.T_bt, T.T_tb, nk, nao, nao, ns, NQ, mesh=(3, 3, 3))
    err = np.abs(osy.full_to_ibz(S, z["k_ibz2bz"]) - g["Sigma_tau"]).max()
    assert err < 1e-6      # reference tolerance (solvers_test.cpp:134)
    assert err < 2e-8      # reached: 8.9e-9 (full mesh here, q stars with rotated P0 in the reference)


def _star(sym, g, rotate=True):
    Vp = oi.packed_from_irreducible(g["V"], 27, *_pairs(sym))
    T = transformer(float(g["lam"]), float(g["beta"]), g["tau"])
    if not rotate:
        sym = dict(sym)
        sym["q_k_sym_transform_p0"] = np.stack([np.eye(36)] * 27)
    return osy.gw_sigma_star(g["G_tau"], Vp, T.T_bt, T.T_tb, sym, 27, 2, 2, 2, 36, (3, 3, 3), return_Pw=True)


def test_gw_sym27_fixture_star_of_q():
    """The reference's own algorithm on the reduced mesh: 4 irreducible q, 4 irreducible k."""
    g = np.load(os.path.join(GOLDEN, "gw_sym27.npz"), allow_pickle=False)
    S, _ = _star(_sym(), g)
    err = np.abs(S - g["Sigma_tau"]).max()
    assert err < 1e-10     # reached: 2.6e-11 (the full-mesh evaluation: 8.9e-9)


def test_gw_sym27_star_without_rotation_is_rejected():
    g = np.load(os.path.join(GOLDEN, "gw_sym27.npz"), allow_pickle=False)
    S, _ = _star(_sym(), g, rotate=False)
    assert np.abs(S - g["Sigma_tau"]).max() > 1e-3


def test_q_stars_partition_the_mesh():
    z = _sym()
    stars = [osy.q_star(z["q_bz2ibz"], z["q_ibz2bz"], iq) for iq in range(int(z["inq"]))]
    assert sorted(q for st in stars for q in st) == list(range(27))
    assert [len(st) for st in stars] == [1, 6, 12, 8]      # input.h5:/symmetry/q/stars
    assert stars[1] == [4, 8, 10, 12, 20, 24]
    for iq, st in enumerate(stars):
        assert int(z["q_ibz2bz"][iq]) in st


def test_p_bz_from_ibz_closed_forms():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((5, 5)) + 1j * rng.standard_normal((5, 5))
    x = x + np.conj(x.T)
    assert np.array_equal(osy.p_bz_from_ibz(x, np.eye(5), 1), np.conj(x))
    u, _ = np.linalg.qr(rng.standard_normal((5, 5)) + 1j * rng.standard_normal((5, 5)))
    y = osy.p_bz_from_ibz(x, u, 0)
    assert np.allclose(np.linalg.eigvalsh(y), np.linalg.eigvalsh(x), atol=1e-12)


def test_planted_pair_operation_swap_is_rejected(monkeypatch):
    z = _sym()
    orig = oi.pair_source

    def swapped(*a):
        row, sign, op = orig(*a)
        return row, sign, {"conj": "trans", "trans": "conj"}.get(op, op)

    monkeypatch.setattr(oi, "pair_source", swapped)
    F, ref = _hf_fixture(z, None)
    assert np.abs(F - ref).max() > 1e-6


def test_planted_transform_order_is_rejected():
    F, ref = _hf_fixture(_sym(), None, transform=lambda x, u: np.conj(u.T) @ x @ u)
    assert np.abs(F - ref).max() > 1e-6


def test_time_reversal_closed_form():
    """A time-reversal partner with U = I is the complex conjugate (real AO basis)."""
    rng = np.random.default_rng(0)
    x = rng.standard_normal((1, 2, 2)) + 1j * rng.standard_normal((1, 2, 2))
    y = osy.value_AO(x, 1, np.array([0, 0]), np.array([0]), np.stack([np.eye(2)] * 2), np.array([0, 1]))
    assert np.array_equal(y, np.conj(x[0]))


# ------------------------------------------------------------------ host reader (libgwhip_h5.so)

FAIL = ["0.2.4", "0.2.4b10", "0.3.0", "0.3.0b8", "0.3.1", "0.3.1b10"]
PASS = ["1.0.0", "1.1.0", "1.1.2b"]


def test_version_strings():
    for v in FAIL:
        assert gl.h5_check_version(v) == 0 and not oi.check_version(v), v
    for v in PASS:
        assert gl.h5_check_version(v) == 1 and oi.check_version(v), v
    assert gl.h5_check_version("1.0") == -1
    with pytest.raises(ValueError):
        oi.check_version("1.0")


def test_meta_versions():
    with pytest.raises(gl.OutdatedInput):
        gl.h5_read_meta(os.path.join(H5, "versions", "df_int"))
    with pytest.raises(gl.OutdatedInput):
        gl.h5_read_meta(os.path.join(H5, "versions", "df_int_0.3.0"))
    assert gl.h5_read_meta(os.path.join(H5, "versions", "df_int_1.0.0")) == 196
    assert gl.h5_read_meta(os.path.join(H5, "versions", "df_int_1.1.0")) == 196


def test_read_packed_matches_oracle_expansion():
    z = _sym()
    g = np.load(os.path.join(GOLDEN, "gw_sym27.npz"), allow_pickle=False)
    V = gl.h5_read_packed(DF, INPUT, 27, 2, 36)
    ref = oi.packed_from_irreducible(g["V"], 27, *_pairs(z))
    assert np.array_equal(V, ref)


def test_read_k_symmetry_and_maps():
    z = _sym()
    s = gl.h5_read_k_symmetry(INPUT, 27, 2)
    assert s["ink"] == int(z["ink"]) == 4
    assert np.array_equal(s["bz2ibz"], z["k_bz2ibz"]) and np.array_equal(s["ibz2bz"], z["k_ibz2bz"])
    assert np.array_equal(s["U"], z["k_k_sym_transform_ao"]) and np.array_equal(s["tr_conj"], z["k_tr_conj"])
    kp, km = gl.h5_kq_maps(s["mesh_scaled"])
    okp, okm = osy.kq_maps(z["k_mesh_scaled"])
    assert np.array_equal(kp, okp) and np.array_equal(km, okm)
    for q in range(27):
        for k in range(27):
            assert kp[q, k] == go.k_add(k, q, (3, 3, 3), 1) and km[q, k] == go.k_add(k, q, (3, 3, 3), -1)


def test_reader_errors():
    with pytest.raises(gl.GwhipError) as e:
        gl.h5_read_packed(os.path.join(H5, "no_such_dir"), INPUT, 27, 2, 36)
    assert e.value.status == gl.ERR_IO
    with pytest.raises(gl.GwhipError) as e:
        gl.h5_read_packed(DF, INPUT, 27, 3, 36)       # wrong nao
    assert e.value.status == gl.ERR_IO
    with pytest.raises(gl.GwhipError) as e:
        gl.h5_read_packed(DF, INPUT, 26, 2, 36)       # pair lists of another mesh
    assert e.value.status == gl.ERR_IO


# ------------------------------------------------------------------ q -> 0 extrapolation (oracle/q0.py)

def _gwc_oracle(plant=None):
    from oracle import q0
    c = np.load(os.path.join(GOLDEN, "gwc_sym27.npz"), allow_pickle=False)
    Vp = oi.packed_from_irreducible(c["V"], 27, c["pairs_conj_pairs_list"], c["pairs_trans_pairs_list"],
                                    c["pairs_kpair_irre_list"])
    G = osy.unfold(c["G_tau"], c["k_bz2ibz"], c["k_ibz2bz"], c["k_k_sym_transform_ao"], c["k_tr_conj"])
    T = transformer(float(c["lam"]), float(c["beta"]), c["tau"])
    S, Pw = go.gw_sigma(G, Vp, T.T_bt, T.T_tb, 27, 2, 2, 2, 36, mesh=(3, 3, 3), return_Pw=True)
    Si = osy.full_to_ibz(S, c["k_ibz2bz"])
    eps = q0.eps_inv_wq(Pw, c["AqQ"], c["q_ibz2bz"])
    if plant == "imag_part":
        eps = eps.imag + 1j * eps.real
    et = q0.eps_q0_tau(eps, c["q_abs"], T.T_tb)
    if plant == "no_correction":
        return Si, c["Sigma_tau"]
    S_ibz = osy.full_to_ibz(c["S_k"], c["k_ibz2bz"])
    return q0.apply_q0_correction(Si, c["G_tau"], S_ibz, et, float(c["q0_madelung"])), c["Sigma_tau"]


def test_gw_c_fixture_with_extrapolation():
    """GW_C section (solvers_test.cpp:390): time-reversal wedge (14 of 27 k) and EXTRAPOLATE."""
    S, ref = _gwc_oracle()
    err = np.abs(S - ref).max()
    assert err < 1e-6      # reference tolerance
    assert err < 1e-10     # reached: 2.8e-11


def test_gw_c_fixture_star_of_q_with_extrapolation():
    """GW_C through the star-of-q evaluation: 14 irreducible q, P-tilde conjugated on the
    time-reversal partners; eps from the irreducible q's P-tilde(i Omega) directly."""
    from oracle import q0
    c = np.load(os.path.join(GOLDEN, "gwc_sym27.npz"), allow_pickle=False)
    sym = {k: c[k] for k in c.files}
    S, Pw = _star(sym, c)
    eps = q0.eps_inv_wq(Pw, c["AqQ"], np.arange(len(c["q_ibz2bz"])))
    et = q0.eps_q0_tau(eps, c["q_abs"], transformer(float(c["lam"]), float(c["beta"]), c["tau"]).T_tb)
    S = q0.apply_q0_correction(S, c["G_tau"], osy.full_to_ibz(c["S_k"], c["k_ibz2bz"]), et, float(c["q0_madelung"]))
    assert np.abs(S - c["Sigma_tau"]).max() < 1e-10     # reached: 2.8e-11


@pytest.mark.parametrize("plant", ["no_correction", "imag_part"])
def test_gw_c_planted_mistakes(plant):
    S, ref = _gwc_oracle(plant)
    assert np.abs(S - ref).max() > 1e-6


def test_q0_fit_closed_form():
    """The extrapolation of an exact quadratic in |q| returns its constant term; duplicates and
    |q| = 0 or > 1 are ignored."""
    from oracle import q0
    qa = np.array([0.0, 0.3, 0.3, 0.5, 0.7, 1.2])
    y = 0.9 - 0.4 * qa + 0.25 * qa ** 2
    y[2] = 99.0          # a repeated |q| value: only its first occurrence is used
    y[5] = -50.0         # beyond q_max
    assert q0.filter_q_abs(qa) == [1, 3, 4]
    assert abs(q0.extrapolate_q0(y.astype(complex), qa) - 0.9) < 1e-12
