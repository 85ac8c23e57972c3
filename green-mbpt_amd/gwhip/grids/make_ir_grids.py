"""Regenerate the committed IR grid files ``ir_1e3.npz`` and ``ir_1e4.npz``.

    python green-mbpt_amd/gwhip/grids/make_ir_grids.py

Each file holds the dimensionless basis (segment edges, piecewise-Legendre
coefficients, singular values), the x-nodes of the fermionic tau grid (roots of
u_L plus the end points) and the bosonic / fermionic Matsubara sampling indices.
``gwhip.irgrid.Transformer.from_grid(data, beta)`` turns them into transform
matrices for a given beta.  The cutoff eps = 1e-14 keeps every basis function
above the float64 noise floor of the SVE (L = 68 at Lambda = 1e3, L = 98 at 1e4).
"""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))

from gwhip import irgrid  # noqa: E402


def main():
    for lam, name in ((1e3, "ir_1e3.npz"), (1e4, "ir_1e4.npz")):
        data = irgrid.build_grid(lam, eps=1e-14)
        irgrid.save_grid(data, os.path.join(irgrid.GRID_DIR, name))
        print(name, "L =", int(data["L"]), "nts =", len(data["x_nodes"]),
              "nw_b =", len(data["n_bose"]), "nw =", len(data["n_fermi"]))


if __name__ == "__main__":
    main()
