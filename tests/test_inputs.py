"""The seeded input generator (test infrastructure): the per-pair V generator used for full-size
checks reproduces the packed store block for block -- CPU only."""
import numpy as np

from gwhip import inputs


def test_lazy_pairs_equal_packed_store():
    V = inputs.generate_V(3, 9, 4, 3, 0.2)
    L = inputs.LazyPackedV(3, 9, 4, 3, 0.2)
    assert len(L) == V.shape[0]
    for p in range(len(L)):
        assert np.array_equal(L[p], V[p])
    for k in range(9):  # diagonal pairs are Hermitian per Q
        d = L[k * (k + 1) // 2 + k]
        assert np.array_equal(d, np.conj(np.transpose(d, (0, 2, 1))))
