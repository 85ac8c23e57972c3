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
* ``HF_orth_invariance/none/data.h5:/iter1/G_tau/mesh`` -- the 110-node IR tau mesh
  (Lambda = 1e4, beta = 100) that the ``ir/1e4.h5`` grid used by all these fixtures has.

Only data (inputs and expected outputs) is written; nothing from the reference's code.
"""
import os
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
    print("wrote", os.listdir(OUT))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test/data")
