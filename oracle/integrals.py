"""CPU restatement of the density-fitted integral reader -- TEST INFRASTRUCTURE ONLY.

Restates ``df_integral_t`` (``src/green/mbpt/df_integral_t.h:40-187``): which stored chunk row
holds V(k1, k2), and the symmetry operations that turn that row into V(k1, k2).  Only ``tests/``
imports it; the library's loader never calls it.

Storage (``df_integral_t.h:64-92``): k pairs are indexed idx = k1(k1+1)/2 + k2 (k1 >= k2).  Of the
nk(nk+1)/2 pairs only the irreducible ones are stored, in the order of ``kpair_irre_list``, in
chunks of ``chunk_size`` rows: file ``VQ_{c}.h5`` with dataset ``/{c}`` [chunk, NQ, nao, 2 nao]
float64 (complex interleaved), c = first row of the chunk.  A pair that is not stored points to
its partner through ``conj_pairs_list`` (stored V conjugated) or ``trans_pairs_list`` (stored V
transposed per Q); the lists are the identity for stored pairs.
"""
from __future__ import annotations

import numpy as np


def pair_index(k1: int, k2: int) -> int:
    """``df_integral_t.h:68``."""
    return k1 * (k1 + 1) // 2 + k2 if k1 >= k2 else k2 * (k2 + 1) // 2 + k1


def irre_pos(kpair_irre_list: np.ndarray, idx: int) -> int:
    """green-symmetry ``irre_pos_kpair``: the row of pair ``idx`` in the stored (irreducible) list."""
    pos = np.flatnonzero(np.asarray(kpair_irre_list) == idx)
    if pos.size != 1:
        raise KeyError(f"k pair {idx} is not in the irreducible list")
    return int(pos[0])


def pair_source(k1: int, k2: int, conj_list, trans_list, irre_list):
    """``read_integrals`` / ``wrap`` / ``v_type`` (``df_integral_t.h:64-84,146-160,181-192``):
    (stored row, sign, op) with sign -1 for k1 < k2 (adjoint per Q) and op in
    {"direct", "conj", "trans"}."""
    idx = pair_index(k1, k2)
    op = "direct"
    if conj_list[idx] != idx:
        idx, op = int(conj_list[idx]), "conj"
    elif trans_list[idx] != idx:
        idx, op = int(trans_list[idx]), "trans"
    return irre_pos(irre_list, idx), (1 if k1 >= k2 else -1), op


def v_k1k2(V_irre: np.ndarray, k1: int, k2: int, conj_list, trans_list, irre_list) -> np.ndarray:
    """``symmetrize`` (``df_integral_t.h:162-187``): V^Q_ij(k1, k2) [NQ, nao, nao].

    V_irre: every stored row in order [n_irre, NQ, nao, nao] (the chunks concatenated)."""
    row, sign, op = pair_source(k1, k2, conj_list, trans_list, irre_list)
    v = V_irre[row]
    if sign < 0:
        v = np.conj(np.swapaxes(v, -1, -2))
    if op == "conj":
        v = np.conj(v)
    elif op == "trans":
        v = np.swapaxes(v, -1, -2)
    return np.ascontiguousarray(v)


def packed_from_irreducible(V_irre, nk, conj_list, trans_list, irre_list) -> np.ndarray:
    """Every pair k1 >= k2 in the packed order [nk(nk+1)/2, NQ, nao, nao] (what the device store
    expands); each row is ``v_k1k2(k1, k2)``."""
    npair = nk * (nk + 1) // 2
    out = np.empty((npair,) + V_irre.shape[1:], np.complex128)
    for k1 in range(nk):
        for k2 in range(k1 + 1):
            out[pair_index(k1, k2)] = v_k1k2(V_irre, k1, k2, conj_list, trans_list, irre_list)
    return out


def check_version(v: str, ref: str = "1.0.0") -> bool:
    """``CheckVersion`` (``common_defs.h:191-212``): True when v >= ref on major.minor.patch; a
    suffix after the patch number is ignored.  Raises ValueError on a malformed string."""
    def parse(s):
        import re
        m = re.match(r"^\s*([+-]?\d+)\.([+-]?\d+)\.([+-]?\d+)", s)
        if not m:
            raise ValueError("Version string format is incorrect. Expected format: major.minor.patch[suffix]")
        return tuple(int(x) for x in m.groups())
    a, b = parse(v), parse(ref)
    return a >= b
