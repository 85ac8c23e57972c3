"""Convert subsets of the reference's own GW test fixtures into small committed vectors.

Run once in a container that has the reference checkout and h5py (the image's
``/opt/conda/bin/python3.9`` has h5py; the main interpreter does not):

    /opt/conda/bin/python3.9 tests/golden/extract_reference_fixtures.py /root/reference/test/data

Sources (reference paths relative to its ``test/data``):

* ``GW_X2C/data.h5``      -- G_tau and result/Sigma_tau used by the ``GW_X2C`` section
  (``test/solvers_test.cpp:391``): nk=1, X2C, nao=10, nso=20, ns=1, NQ=154, beta=100.
* ``GW_X2C/df_hf_int/VQ_0.h5`` -- the one k-pair chunk of V^Q_ij, stored as
  [chunk, NQ, nao, 2*nao] float64 (``df_integral_t.h:86-92`` reads it as complex).
* ``embedding/data.h5`` + ``embedding/dc_int.0`` -- the ``SEET DC TEST`` inputs
  (``test/solvers_test.cpp:602-672``): nk=1, ns=2, nao=nso=2, NQ=36 of which the
  file holds 3 rows (the reader zero-fills the rest, ``df_integral_t.h:87``).
* ``HF_X2C/data.h5`` + ``HF_X2C/input.h5`` + ``HF_X2C/df_hf_int/VQ_0.h5`` -- the ``HF_X2C``
  section (``test/solvers_test.cpp:388``, ``solve_hf`` at ``:27-79``): G(beta), the overlap
  S-k, the Madelung constant and V^Q_ij for nk=1, X2C, nao=10, NQ=154, with the expected
  result/Sigma1 (the X2C Hartree-Fock Fock matrix).
* ``GW/input.h5`` (identical to ``HF/input.h5``) -- the nk=27 (3x3x3) H2 system with space-group
  symmetry: ns=2, nao=nso=2, NQ=36, ink=4 irreducible k points, 196 of 378 k pairs stored.
  Written to ``sym27.npz``: params, ``symmetry/k`` (bz2ibz, ibz2bz, k_sym_transform_ao,
  tr_conj, mesh_scaled), ``symmetry/q`` (inq, bz2ibz, ibz2bz, tr_conj and the auxiliary-basis
  matrices k_sym_transform_p0), ``symmetry/pairs`` (conj/trans pair lists, kpair_irre_list,
  kpair_idx), HF/S-k and HF/madelung.
* ``HF/data.h5`` + ``HF/df_hf_int/VQ_0.h5`` -- the ``HF`` section (``solvers_test.cpp:387``):
  G_tau (irreducible k), result/Sigma1 -> ``hf_sym27.npz``.
* ``GW/data.h5`` + ``GW/df_int/VQ_0.h5`` -- the ``GW`` section (``solvers_test.cpp:389``):
  G_tau, result/Sigma_tau (irreducible k) -> ``gw_sym27.npz``.
* The integral files themselves, byte for byte, for the HDF5 loader: ``GW/df_int/{meta,VQ_0}.h5``
  -> ``h5/gw_sym27_df/``; ``Input/df_int*/meta.h5`` (no version attribute, 0.3.0, 1.0.0, 1.1.0;
  ``solvers_test.cpp:470-495``) -> ``h5/versions/<dir>/meta.h5``; and the ``params``,
  ``symmetry/k`` and ``symmetry/pairs`` groups and the ``symmetry/q`` tables of ``GW/input.h5``
  -> ``h5/sym27_input.h5``.
* ``GW/input_tr_symm.h5`` + ``GW/df_hf_int/{VQ_0,AqQ}.h5`` + ``GW/data_c.h5`` -- the ``GW_C``
  section (``solvers_test.cpp:390``, q0_treatment EXTRAPOLATE): the same H2 system with 14
  irreducible k (time reversal on 13 of the 27 points), the q -> 0 data (AqQ [inq][NQ], q_abs,
  madelung) and G_tau / result/Sigma_tau -> ``gwc_sym27.npz``; the integral directory byte for byte
  -> ``h5/gw_c_df/``; ``params``, ``symmetry/k``, ``symmetry/pairs`` and the small ``symmetry/q``
  tables -> ``h5/sym27_tr_input.h5``.
* ``HF_orth_invariance/{none,lowdin,mo,natural}/{input,data}.h5`` (``solvers_test.cpp:296-452``): the
  stored first iteration of an SC run (nk=1, ns=2, nao=10, nel=2): HF/H-k, HF/S-k, iter1/Sigma1,
  iter1/Selfenergy, iter1/mu and the Dyson output iter1/G_tau -> ``dyson_orth.npz``.
* ``HF_orth_invariance/none/data.h5:/iter1/G_tau/mesh`` -- the 110-node IR tau mesh
  (Lambda = 1e4, beta = 100) that the ``ir/1e4.h5`` grid used by all these fixtures has.

Only data (inputs and expected outputs) is written; nothing from the reference's code.
"""
import os
import shutil
import sys

import h5py
import numpy as np

OUT = os.path.dirname(os.path.abspath(__file__))


def _v_chunk(path):
    with h5py.File(path, "r") as f:
        raw = f["0"][()]
    return np.ascontiguousarray(raw).view(np.complex128)   # [chunk, NQ, nao, nao]


def main(data_dir):
    with h5py.File(os.path.join(data_dir, "HF_orth_invariance/none/data.h5"), "r") as f:
        mesh = f["iter1/G_tau/mesh"][()]

    with h5py.File(os.path.join(data_dir, "GW_X2C/data.h5"), "r") as f:
        G = f["G_tau"][()]
        S = f["result/Sigma_tau"][()]
    with h5py.File(os.path.join(data_dir, "GW_X2C/input.h5"), "r") as f:
        p = {k: int(f["params/" + k][()]) for k in ("nao", "nso", "ns", "NQ", "nk")}
    V = _v_chunk(os.path.join(data_dir, "GW_X2C/df_hf_int/VQ_0.h5"))
    np.savez(os.path.join(OUT, "gw_x2c.npz"), G_tau=G, Sigma_tau=S, V=V, tau=mesh, beta=100.0,
             lam=1e4, **p)

    with h5py.File(os.path.join(data_dir, "embedding/data.h5"), "r") as f:
        g = f["g_tau"][()]
        s = f["sigma_tau"][()]
    V3 = _v_chunk(os.path.join(data_dir, "embedding/dc_int.0/VQ_0.h5"))
    with h5py.File(os.path.join(data_dir, "embedding/dc_int.0/dummy.h5"), "r") as f:
        NQ = int(f["params/NQ"][()])
        nao = int(f["params/nao"][()])
    np.savez(os.path.join(OUT, "seet_dc.npz"), G_tau=g, Sigma_tau=s, V_rows=V3, tau=mesh,
             beta=100.0, lam=1e4, nao=nao, nso=nao, ns=2, NQ=NQ, nk=1)
    with h5py.File(os.path.join(data_dir, "HF_X2C/data.h5"), "r") as f:
        G_beta = f["G_tau"][-1]
        S1 = f["result/Sigma1"][()]
    with h5py.File(os.path.join(data_dir, "HF_X2C/input.h5"), "r") as f:
        p = {k: int(f["params/" + k][()]) for k in ("nao", "nso", "ns", "NQ", "nk")}
        S_k = np.ascontiguousarray(f["HF/S-k"][()]).view(np.complex128)[..., 0]
        madelung = float(f["HF/madelung"][()])
    V = _v_chunk(os.path.join(data_dir, "HF_X2C/df_hf_int/VQ_0.h5"))
    np.savez(os.path.join(OUT, "hf_x2c.npz"), G_beta=G_beta, Sigma1=S1, S_k=S_k, madelung=madelung, V=V, **p)
    sym_in = os.path.join(data_dir, "GW/input.h5")
    with h5py.File(sym_in, "r") as f:
        sym = {k: int(f["params/" + k][()]) for k in ("nao", "nso", "ns", "NQ", "nk")}
        sym["ink"] = int(f["symmetry/k/ink"][()])
        for k in ("bz2ibz", "ibz2bz", "k_sym_transform_ao", "tr_conj", "mesh_scaled"):
            sym["k_" + k] = f["symmetry/k/" + k][()]
        for k in ("conj_pairs_list", "trans_pairs_list", "kpair_irre_list", "kpair_idx", "num_kpair_stored"):
            sym["pairs_" + k] = f["symmetry/pairs/" + k][()]
        sym["inq"] = int(f["symmetry/q/inq"][()])
        for k in ("bz2ibz", "ibz2bz", "tr_conj", "k_sym_transform_p0"):
            sym["q_" + k] = f["symmetry/q/" + k][()]
        sym["S_k"] = np.ascontiguousarray(f["HF/S-k"][()]).view(np.complex128)[..., 0]
        sym["madelung"] = float(f["HF/madelung"][()])
    np.savez(os.path.join(OUT, "sym27.npz"), **sym)
    with h5py.File(os.path.join(data_dir, "HF/data.h5"), "r") as f:
        hf = dict(G_tau=f["G_tau"][()], Sigma1=f["result/Sigma1"][()])
    np.savez(os.path.join(OUT, "hf_sym27.npz"), V=_v_chunk(os.path.join(data_dir, "HF/df_hf_int/VQ_0.h5")), **hf)
    with h5py.File(os.path.join(data_dir, "GW/data.h5"), "r") as f:
        gw = dict(G_tau=f["G_tau"][()], Sigma_tau=f["result/Sigma_tau"][()])
    np.savez(os.path.join(OUT, "gw_sym27.npz"), V=_v_chunk(os.path.join(data_dir, "GW/df_int/VQ_0.h5")), tau=mesh,
             beta=100.0, lam=1e4, **gw)

    # integral files for the HDF5 loader (data files of the reference's tests, copied unchanged)
    h5dir = os.path.join(OUT, "h5")
    os.makedirs(os.path.join(h5dir, "gw_sym27_df"), exist_ok=True)
    for name in ("meta.h5", "VQ_0.h5"):
        shutil.copyfile(os.path.join(data_dir, "GW/df_int", name), os.path.join(h5dir, "gw_sym27_df", name))
    for d in ("df_int", "df_int_0.3.0", "df_int_1.0.0", "df_int_1.1.0"):
        os.makedirs(os.path.join(h5dir, "versions", d), exist_ok=True)
        shutil.copyfile(os.path.join(data_dir, "Input", d, "meta.h5"), os.path.join(h5dir, "versions", d, "meta.h5"))
    with h5py.File(sym_in, "r") as src, h5py.File(os.path.join(h5dir, "sym27_input.h5"), "w") as dst:
        for a, v in src.attrs.items():
            dst.attrs[a] = v
        src.copy(src["params"], dst, name="params")
        dst.create_group("symmetry")
        for g in ("k", "pairs"):
            src.copy(src["symmetry/" + g], dst["symmetry"], name=g)
        q = dst["symmetry"].create_group("q")
        for k in ("bz2ibz", "ibz2bz", "inq", "tr_conj", "k_sym_transform_p0", "mesh_scaled", "nq"):
            q[k] = src["symmetry/q/" + k][()]
    tr_in = os.path.join(data_dir, "GW/input_tr_symm.h5")
    c = {}
    with h5py.File(tr_in, "r") as f:
        c["ink"] = int(f["symmetry/k/ink"][()])
        c["inq"] = int(f["symmetry/q/inq"][()])
        for k in ("bz2ibz", "ibz2bz", "k_sym_transform_ao", "tr_conj", "mesh_scaled"):
            c["k_" + k] = f["symmetry/k/" + k][()]
        for k in ("bz2ibz", "ibz2bz", "tr_conj", "k_sym_transform_p0"):
            c["q_" + k] = f["symmetry/q/" + k][()]
        for k in ("conj_pairs_list", "trans_pairs_list", "kpair_irre_list"):
            c["pairs_" + k] = f["symmetry/pairs/" + k][()]
        c["S_k"] = np.ascontiguousarray(f["HF/S-k"][()]).view(np.complex128)[..., 0]
    with h5py.File(os.path.join(data_dir, "GW/df_hf_int/AqQ.h5"), "r") as f:
        c["AqQ"] = f["AqQ"][()]
        c["q_abs"] = f["q_abs"][()]
        c["q0_madelung"] = float(f["madelung"][()])
    with h5py.File(os.path.join(data_dir, "GW/data_c.h5"), "r") as f:
        c["G_tau"] = f["G_tau"][()]
        c["Sigma_tau"] = f["result/Sigma_tau"][()]
    np.savez(os.path.join(OUT, "gwc_sym27.npz"), V=_v_chunk(os.path.join(data_dir, "GW/df_hf_int/VQ_0.h5")), tau=mesh,
             beta=100.0, lam=1e4, **c)
    os.makedirs(os.path.join(h5dir, "gw_c_df"), exist_ok=True)
    for name in ("meta.h5", "VQ_0.h5", "AqQ.h5"):
        shutil.copyfile(os.path.join(data_dir, "GW/df_hf_int", name), os.path.join(h5dir, "gw_c_df", name))
    with h5py.File(tr_in, "r") as src, h5py.File(os.path.join(h5dir, "sym27_tr_input.h5"), "w") as dst:
        for a, v in src.attrs.items():
            dst.attrs[a] = v
        src.copy(src["params"], dst, name="params")
        dst.create_group("symmetry")
        for g in ("k", "pairs"):
            src.copy(src["symmetry/" + g], dst["symmetry"], name=g)
        q = dst["symmetry"].create_group("q")
        for k in ("bz2ibz", "ibz2bz", "inq", "tr_conj", "k_sym_transform_p0", "mesh", "mesh_scaled", "nq",
                  "weight_ibz"):
            q[k] = src["symmetry/q/" + k][()]
    orth = {}
    for mode in ("none", "lowdin", "mo", "natural"):
        base = os.path.join(data_dir, "HF_orth_invariance", mode)
        with h5py.File(os.path.join(base, "input.h5"), "r") as f:
            orth[mode + "_H"] = np.ascontiguousarray(f["HF/H-k"][()]).view(np.complex128)[..., 0]
            orth[mode + "_S"] = np.ascontiguousarray(f["HF/S-k"][()]).view(np.complex128)[..., 0]
            orth[mode + "_nel"] = int(f["params/nel_cell"][()])
        with h5py.File(os.path.join(base, "data.h5"), "r") as f:
            orth[mode + "_Sigma1"] = f["iter1/Sigma1"][()]
            orth[mode + "_Sigma_tau"] = f["iter1/Selfenergy/data"][()]
            orth[mode + "_G_tau"] = f["iter1/G_tau/data"][()]
            orth[mode + "_mu"] = float(f["iter1/mu"][()])
    np.savez_compressed(os.path.join(OUT, "dyson_orth.npz"), tau=mesh, beta=100.0, lam=1e4, **orth)
    extract_orth_energies(data_dir)
    print("wrote", os.listdir(OUT))


def extract_orth_energies(data_dir: str):
    """The energies the reference stored for the first iteration of each HF_orth_invariance run
    (iter1/Energy_1b, iter1/Energy_HF = E_HF + E_nuc, iter1/Energy_2b; HF/Energy_nuc of input.h5):
    compute_energy's outputs (src/common_utils.cpp:5-55, written by dyson.cpp:319-339) for the
    G_tau, Sigma1 and Sigma_tau of ``dyson_orth.npz`` -> ``orth_energies.npz``."""
    out = {}
    for mode in ("none", "lowdin", "mo", "natural"):
        base = os.path.join(data_dir, "HF_orth_invariance", mode)
        with h5py.File(os.path.join(base, "input.h5"), "r") as f:
            out[mode + "_E_nuc"] = float(f["HF/Energy_nuc"][()])
        with h5py.File(os.path.join(base, "data.h5"), "r") as f:
            for key in ("Energy_1b", "Energy_HF", "Energy_2b"):
                out[mode + "_" + key] = float(f["iter1/" + key][()])
    np.savez(os.path.join(OUT, "orth_energies.npz"), **out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test/data")
