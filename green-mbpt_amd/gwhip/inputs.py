"""Seeded synthetic inputs for the GW self-energy path.

* G(tau, s, k) = -U_k diag(f(eps_k, tau)) U_k^H with f = e^{-eps tau}/(1 + e^{-beta eps})
  (evaluated as e^{eps (beta - tau)}/(1 + e^{beta eps}) for eps < 0), U_k Haar-random
  (QR of a complex Gaussian, phases fixed by R's diagonal), eps ~ U[-1, 1] Ha, mu = 0.
  With ns = 2 the two spins get F_k -/+ 0.05 Ha (same U_k, shifted eigenvalues).
* V^Q_ij(k1, k2) for the packed pairs k1 >= k2 (pair index k1 (k1+1)/2 + k2, the store
  of ``df_integral_t``): real and imaginary parts ~ N(0, sigma^2); the k1 == k2 blocks
  are Hermitised per Q so that V^Q_ij(k1, k2) = conj(V^Q_ji(k2, k1)) holds for all pairs.

Random numbers come from a counter-based generator (SplitMix64 of seed, stream and
element index, Box-Muller), so any element can be produced independently; the HIP
library implements the same counters (``gwhip_generate_V``), which makes the device
copy of V equal to this one up to the last bits of log/cos.
"""
from __future__ import annotations

import math

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
STREAM_V = 0x5611
STREAM_U = 0x7A11
STREAM_EPS = 0x3E95


def splitmix64(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def stream_key(seed: int, stream: int) -> np.uint64:
    return splitmix64(np.uint64(seed) ^ splitmix64(np.uint64(stream)))


def uniform(seed: int, stream: int, idx: np.ndarray) -> np.ndarray:
    """Uniform doubles in (0, 1) for counters idx."""
    key = stream_key(seed, stream)
    with np.errstate(over="ignore"):
        u = splitmix64(key + np.asarray(idx, dtype=np.uint64))
    return ((u >> np.uint64(11)).astype(np.float64) + 0.5) * (2.0 ** -53)


def normal_pairs(seed: int, stream: int, idx: np.ndarray):
    """Two independent N(0,1) per counter idx (Box-Muller on counters 2 idx, 2 idx + 1)."""
    idx = np.asarray(idx, dtype=np.uint64)
    u1 = uniform(seed, stream, np.uint64(2) * idx)
    u2 = uniform(seed, stream, np.uint64(2) * idx + np.uint64(1))
    r = np.sqrt(-2.0 * np.log(u1))
    return r * np.cos(2.0 * np.pi * u2), r * np.sin(2.0 * np.pi * u2)


def pair_count(nk: int) -> int:
    return nk * (nk + 1) // 2


def generate_V(seed: int, nk: int, NQ: int, nao: int, sigma: float = 0.1) -> np.ndarray:
    """Packed V [nk(nk+1)/2, NQ, nao, nao] complex128."""
    npair = pair_count(nk)
    n = npair * NQ * nao * nao
    re, im = normal_pairs(seed, STREAM_V, np.arange(n, dtype=np.uint64))
    V = (sigma * (re + 1j * im)).reshape(npair, NQ, nao, nao)
    for k in range(nk):
        p = k * (k + 1) // 2 + k
        V[p] = 0.5 * (V[p] + np.conj(np.transpose(V[p], (0, 2, 1))))
    return V


def generate_V_pair(seed: int, pair: int, NQ: int, nao: int, sigma: float = 0.1) -> np.ndarray:
    """One block [NQ, nao, nao] of the packed V of ``generate_V`` (same counters), without the rest."""
    n = NQ * nao * nao
    idx = np.uint64(pair) * np.uint64(n) + np.arange(n, dtype=np.uint64)
    re, im = normal_pairs(seed, STREAM_V, idx)
    v = (sigma * (re + 1j * im)).reshape(NQ, nao, nao)
    k = (math.isqrt(8 * int(pair) + 1) - 1) // 2       # pair = k1 (k1 + 1)/2 + k2 with k2 <= k1
    if pair - k * (k + 1) // 2 == k:                   # k1 == k2: Hermitised per Q
        v = 0.5 * (v + np.conj(np.transpose(v, (0, 2, 1))))
    return v


class LazyPackedV:
    """The packed V of ``generate_V`` indexed by pair, each block generated when first read (for
    meshes whose whole packed store would not fit host memory); keeps the blocks it has made."""

    def __init__(self, seed: int, nk: int, NQ: int, nao: int, sigma: float = 0.1):
        self.seed, self.nk, self.NQ, self.nao, self.sigma = seed, nk, NQ, nao, sigma
        self._cache = {}

    def __len__(self):
        return pair_count(self.nk)

    def __getitem__(self, pair: int) -> np.ndarray:
        pair = int(pair)
        if not 0 <= pair < len(self):
            raise IndexError(pair)
        if pair not in self._cache:
            self._cache[pair] = generate_V_pair(self.seed, pair, self.NQ, self.nao, self.sigma)
        return self._cache[pair]


def fock_eigensystem(seed: int, nk: int, nso: int):
    """(U [nk, nso, nso], eps [nk, nso]) -- Haar unitaries and U[-1, 1] eigenvalues."""
    n = nk * nso * nso
    re, im = normal_pairs(seed, STREAM_U, np.arange(n, dtype=np.uint64))
    Z = (re + 1j * im).reshape(nk, nso, nso) / np.sqrt(2.0)
    U = np.empty_like(Z)
    for k in range(nk):
        Qm, Rm = np.linalg.qr(Z[k])
        d = np.diagonal(Rm)
        U[k] = Qm * (d / np.abs(d))[None, :]
    eps = 2.0 * uniform(seed, STREAM_EPS, np.arange(nk * nso, dtype=np.uint64)) - 1.0
    return U, eps.reshape(nk, nso)


def fermi_tau(eps: np.ndarray, tau: np.ndarray, beta: float) -> np.ndarray:
    """e^{-eps tau}/(1 + e^{-beta eps}) without overflow; shape [len(tau)] + eps.shape."""
    e = np.asarray(eps)[None, ...]
    t = np.asarray(tau).reshape((-1,) + (1,) * np.ndim(eps))
    pos = np.exp(-np.abs(e) * t) / (1.0 + np.exp(-beta * np.abs(e)))
    neg = np.exp(-np.abs(e) * (beta - t)) / (1.0 + np.exp(-beta * np.abs(e)))
    return np.where(e >= 0, pos, neg)


def generate_G(seed: int, nk: int, nso: int, ns: int, tau: np.ndarray, beta: float,
               spin_split: float = 0.05) -> np.ndarray:
    """G [nts, ns, nk, nso, nso] complex128 of a non-interacting Fock operator."""
    U, eps = fock_eigensystem(seed, nk, nso)
    nts = len(tau)
    G = np.empty((nts, ns, nk, nso, nso), dtype=np.complex128)
    for s in range(ns):
        shift = 0.0 if ns == 1 else (-spin_split if s == 0 else spin_split)
        f = fermi_tau(eps + shift, tau, beta)                       # [nts, nk, nso]
        G[:, s] = -np.einsum("kia,tka,kja->tkij", U, f, np.conj(U))
    return G
