"""Intermediate-representation (IR) imaginary-time grids and tau <-> Matsubara transforms.

The reference reaches these through green-grids' ``transformer_t`` (not vendored in
green-mbpt; call sites ``src/gw_cpu_kernel.cpp:234`` ``tau_f_to_w_b`` and ``:264``
``w_b_to_tau_f``, the grid sizes in ``src/green/mbpt/kernels.h:44-45``, and
``src/common_utils.cpp:22,31-32`` for the fermionic transforms used by the
Galitskii-Migdal energy).  This module rebuilds the same kind of object from the
published IR construction (Shinaoka et al., PRB 96, 035147 (2017); Chikano et al.,
CPC 240, 181 (2019); Wallerberger et al., SoftwareX 21, 101266 (2023)):

1. singular-value expansion (SVE) of the logistic kernel
   K(x, y) = exp(-Lambda*y*(x+1)/2) / (1 + exp(-Lambda*y)),  x, y in [-1, 1],
   discretised with composite Gauss-Legendre quadrature on segments graded towards
   the kernel's boundary layers, split into even/odd sectors (K(-x,-y) = K(x,y));
2. basis functions u_l(tau) = sqrt(2/beta) u_l(x), x = 2 tau/beta - 1, stored as
   piecewise Legendre series; their Matsubara transforms
   u_l(i nu) = int_0^beta dtau exp(i nu tau) u_l(tau) are evaluated segment by
   segment with  int_{-1}^{1} exp(i a t) P_k(t) dt = 2 i^k j_k(a);
3. fermionic tau nodes = roots of u_L plus the end points 0 and beta (so
   nts = L + 2, the convention of ``test/solvers_test.cpp:111-114``); Matsubara
   sampling frequencies = sign changes of the highest basis function's transform,
   mirrored to negative frequencies;
4. least-squares transforms between the tau nodes and the Matsubara frequencies.

Fourier convention: f(i nu_n) = int_0^beta dtau exp(i nu_n tau) f(tau); nu_n = 2n pi/beta
(bosonic) or (2n+1) pi/beta (fermionic).

The SVE runs in float64.  Singular values below about 1e-15 s_0 are at the float64
noise floor, so grids are generated with a cutoff where the functions are reliable;
tau nodes can also be supplied externally (``Transformer.from_nodes``), which is how
the tests use the tau mesh stored with the reference's own Sigma fixtures.
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np
from numpy.polynomial.legendre import leggauss, legval
from scipy.special import spherical_jn
from scipy.optimize import brentq

GRID_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "grids")


def _logistic(x: np.ndarray, y: np.ndarray, lam: float) -> np.ndarray:
    """K(x, y) for x, y in [-1, 1], evaluated without overflow."""
    x = np.asarray(x, dtype=np.float64)[:, None]
    y = np.asarray(y, dtype=np.float64)[None, :]
    ay = np.abs(y)
    den = 1.0 + np.exp(-lam * ay)
    # y >= 0: exp(-lam y (x+1)/2);  y < 0: exp(lam |y| (x+1)/2 - lam |y|) = exp(-lam |y| (1-x)/2)
    num = np.where(y >= 0, np.exp(-lam * ay * (x + 1) / 2), np.exp(-lam * ay * (1 - x) / 2))
    return num / den


def _graded_edges(lam: float, finest: float = 0.05) -> np.ndarray:
    """Segment edges on [0, 1] refined geometrically towards 1 down to ~finest/lam."""
    nlev = int(np.ceil(np.log2(lam / finest)))
    inner = [1.0 - 2.0 ** (-j) for j in range(1, nlev + 1)]
    return np.array(sorted(set([0.0, 0.25] + inner + [1.0])))


def _gauss_nodes(edges: np.ndarray, ng: int):
    t, w = leggauss(ng)
    a, b = edges[:-1, None], edges[1:, None]
    x = ((b - a) / 2 * t + (a + b) / 2).ravel()
    wx = ((b - a) / 2 * w).ravel()
    return x, wx


@dataclass
class Basis:
    """Piecewise-Legendre IR basis on x in [-1, 1] for one Lambda.

    ``coef[l, s, k]`` is the k-th Legendre coefficient of u_l on the positive
    segment s = [edges[s], edges[s+1]]; u_l(-x) = (-1)^l u_l(x).  Functions are
    normalised so that int_{-1}^{1} u_l(x)^2 dx = 1.
    """

    lam: float
    edges: np.ndarray     # positive-half segment edges, edges[0] = 0, edges[-1] = 1
    coef: np.ndarray      # [nfunc, nseg, ng]
    s: np.ndarray         # singular values, descending

    @property
    def size(self) -> int:
        return self.coef.shape[0]

    def u(self, x: np.ndarray, lmax: int | None = None) -> np.ndarray:
        """u_l(x) for l < lmax at points x in [-1, 1]; returns [len(x), lmax]."""
        x = np.atleast_1d(np.asarray(x, dtype=np.float64))
        lmax = self.size if lmax is None else lmax
        ax = np.abs(x)
        seg = np.clip(np.searchsorted(self.edges, ax, side="right") - 1, 0, len(self.edges) - 2)
        a, b = self.edges[seg], self.edges[seg + 1]
        t = (2 * ax - a - b) / (b - a)
        out = np.empty((len(x), lmax))
        for l in range(lmax):
            c = self.coef[l, seg, :]                       # [npts, ng]
            vals = np.array([legval(ti, ci) for ti, ci in zip(t, c)])
            par = 1.0 if l % 2 == 0 else -1.0
            out[:, l] = np.where(x >= 0, vals, par * vals)
        return out

    def uhat(self, w: np.ndarray, lmax: int | None = None) -> np.ndarray:
        """Dimensionless Matsubara transforms.

        Returns  e^{i pi w/2} int_{-1}^{1} e^{i pi w x / 2} u_l(x) dx  for integer w
        (w = 2n bosonic, 2n+1 fermionic); the physical transform is sqrt(beta/2) times
        this.  Shape [len(w), lmax].
        """
        w = np.atleast_1d(np.asarray(w, dtype=np.float64))
        lmax = self.size if lmax is None else lmax
        om = np.pi * w / 2.0
        a, b = self.edges[:-1], self.edges[1:]
        h, m = (b - a) / 2, (a + b) / 2                   # [nseg]
        ng = self.coef.shape[2]
        k = np.arange(ng)
        # J[w, s, k] = 2 j_k(om h_s); positive segment integral of P_k(t(x)) e^{i om x}
        arg = om[:, None, None] * h[None, :, None]
        J = 2.0 * spherical_jn(k[None, None, :], arg)
        ik = (1j) ** k
        out = np.empty((len(w), lmax), dtype=np.complex128)
        for l in range(lmax):
            c = self.coef[l]                                 # [nseg, ng]
            # int over [0,1]: sum_s h_s e^{i om m_s} sum_k c_k i^k J
            pos = np.einsum("s,ws,sk,k,wsk->w", h, np.exp(1j * om[:, None] * m[None, :]), c, ik, J)
            # mirrored half: u(-x) = par u(x)  ->  int_{-1}^{0} e^{i om x} u(x) dx = par conj(pos)
            par = 1.0 if l % 2 == 0 else -1.0
            tot = pos + par * np.conj(pos)
            out[:, l] = np.exp(1j * om) * tot
        return out


def sve_logistic(lam: float, nfunc: int | None = None, ng: int = 20, finest: float = 0.05) -> Basis:
    """Float64 SVE of the logistic kernel, even and odd sectors separately."""
    ex = _graded_edges(lam, finest)           # x: graded towards x = 1 (tau -> beta)
    ey = 1.0 - ex[::-1]                       # y: graded towards y = 0 (omega -> 0)
    x, wx = _gauss_nodes(ex, ng)
    y, wy = _gauss_nodes(ey, ng)
    nseg = len(ex) - 1
    tq, wq = leggauss(ng)
    leg = np.array([legval(tq, np.eye(ng)[k]) for k in range(ng)])   # P_k(t_i) [ng, ng]
    proj = (2 * np.arange(ng)[:, None] + 1) / 2.0 * leg * wq[None, :]  # c_k = sum_i proj[k,i] f_i
    funcs, svals = [], []
    for sign in (1.0, -1.0):
        K = _logistic(x, y, lam) + sign * _logistic(x, -y, lam)
        A = np.sqrt(wx)[:, None] * K * np.sqrt(wy)[None, :]
        U, s, _ = np.linalg.svd(A, full_matrices=False)
        ux = U / np.sqrt(wx)[:, None] / np.sqrt(2.0)    # full-domain normalisation
        # fix sign: u_l(1) > 0
        for j in range(ux.shape[1]):
            if ux[-1, j] < 0:
                ux[:, j] *= -1
        vals = ux.T.reshape(-1, nseg, ng)                  # [nf, nseg, ng]
        funcs.append(np.einsum("ki,fsi->fsk", proj, vals))
        svals.append(s)
    # interleave: l even <- even sector, l odd <- odd sector (strict interleaving of s)
    ne = min(len(svals[0]), len(svals[1]))
    n_out = 2 * ne if nfunc is None else nfunc
    coef = np.empty((n_out,) + funcs[0].shape[1:])
    s_out = np.empty(n_out)
    for l in range(n_out):
        coef[l] = funcs[l % 2][l // 2]
        s_out[l] = svals[l % 2][l // 2]
    return Basis(lam=lam, edges=ex, coef=coef, s=s_out)


def basis_size(b: Basis, eps: float) -> int:
    return int(np.sum(b.s / b.s[0] > eps))


def tau_roots(b: Basis, l: int) -> np.ndarray:
    """All roots of u_l on (-1, 1) (l roots), sorted."""
    roots = []
    for si in range(len(b.edges) - 1):
        a, bb = b.edges[si], b.edges[si + 1]
        c = b.coef[l, si]
        f = lambda xx: legval((2 * xx - a - bb) / (bb - a), c)
        xs = np.linspace(a, bb, 400)
        fv = f(xs)
        for j in range(len(xs) - 1):
            if fv[j] == 0.0:
                roots.append(xs[j])
            elif fv[j] * fv[j + 1] < 0:
                roots.append(brentq(f, xs[j], xs[j + 1], xtol=1e-16, rtol=1e-15))
    pos = np.array(sorted(set([r for r in roots if r > 0])))
    par_root = [0.0] if l % 2 == 1 else []
    allr = np.concatenate([-pos[::-1], par_root, pos])
    return allr


def matsubara_sign_changes(b: Basis, l: int, zeta: int, nmax: int) -> np.ndarray:
    """Non-negative Matsubara indices n where the non-vanishing part of uhat_l changes sign.

    For each sign change between consecutive integers n, n+1 the index with the
    smaller magnitude is kept (w = 2n + zeta, zeta = 0 bosonic, 1 fermionic).
    """
    dense = np.arange(0, min(nmax, 4000))
    geo = np.unique(np.round(np.geomspace(4000, max(nmax, 4001), 800)).astype(np.int64))
    grid = np.unique(np.concatenate([dense, geo]))
    vals = b.uhat(2 * grid + zeta, lmax=l + 1)[:, l]
    use_real = np.abs(np.real(vals)).max() > np.abs(np.imag(vals)).max()
    part = np.real(vals) if use_real else np.imag(vals)

    def g(n):
        v = b.uhat(np.array([2 * n + zeta]), lmax=l + 1)[0, l]
        return np.real(v) if use_real else np.imag(v)

    pts = []
    for j in range(len(grid) - 1):
        if part[j] == 0.0:
            pts.append(int(grid[j]))
            continue
        if part[j] * part[j + 1] < 0:
            lo, hi = int(grid[j]), int(grid[j + 1])
            flo = part[j]
            while hi - lo > 1:
                mid = (lo + hi) // 2
                fm = g(mid)
                if fm * flo > 0:
                    lo, flo = mid, fm
                else:
                    hi = mid
            pts.append(lo if abs(flo) < abs(g(hi)) else hi)
    return np.array(sorted(set(pts)), dtype=np.int64)


def _symmetric_freqs(npos: np.ndarray, zeta: int) -> np.ndarray:
    """Mirror non-negative indices n to the set {n} U {-n - zeta}; sorted."""
    neg = -npos - zeta
    return np.array(sorted(set(npos.tolist()) | set(neg.tolist())), dtype=np.int64)


def build_grid(lam: float, eps: float = 1e-14, ng: int = 20) -> dict:
    """Dimensionless grid data for one Lambda (everything needed by ``Transformer``)."""
    b = sve_logistic(lam, ng=ng)
    L = basis_size(b, eps)
    x_int = tau_roots(b, L)
    assert len(x_int) == L, (len(x_int), L)
    nmax = int(20 * lam)
    # sampling frequencies from the first function outside the basis whose parity gives
    # at least L sign changes over all integers (fermionic: even index; bosonic: odd index)
    lf = L if L % 2 == 0 else L + 1
    lb = L if L % 2 == 1 else L + 1
    nf = matsubara_sign_changes(b, lf, 1, nmax)
    nb = matsubara_sign_changes(b, lb, 0, nmax)
    wb = _symmetric_freqs(nb, 0)
    wf = _symmetric_freqs(nf, 1)
    return dict(lam=np.float64(lam), eps=np.float64(eps), L=np.int64(L),
                x_nodes=np.concatenate([[-1.0], x_int, [1.0]]),
                n_bose=wb, n_fermi=wf, s=b.s[:L + 2],
                edges=b.edges, coef=b.coef[:L + 2])


def save_grid(data: dict, path: str) -> None:
    np.savez(path, **data)


def load_grid(lam: float) -> dict:
    name = os.path.join(GRID_DIR, f"ir_{lam:.0e}.npz".replace("+0", "").replace("+", ""))
    with np.load(name, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


class Transformer:
    """tau <-> Matsubara transforms for one (Lambda, beta), dimensioned.

    Mirrors the pieces of green-grids' ``transformer_t`` used on the GW path:
    ``nts`` (fermionic tau nodes including 0 and beta), ``nw_b`` (bosonic Matsubara
    sampling frequencies), ``T_bt`` = tau_f_to_w_b as a [nw_b, nts] matrix,
    ``T_tb`` = w_b_to_tau_f as [nts, nw_b], and the fermionic pair ``T_ft``
    [nw, nts], ``T_tf`` [nts, nw] used for the Galitskii-Migdal energy
    (``src/common_utils.cpp:22,31-36``).
    """

    def __init__(self, basis: Basis, L: int, beta: float, x_nodes: np.ndarray,
                 n_bose: np.ndarray, n_fermi: np.ndarray):
        self.beta = float(beta)
        self.L = int(L)
        self.x_nodes = np.asarray(x_nodes, dtype=np.float64)
        self.tau = self.beta * (self.x_nodes + 1.0) / 2.0
        self.n_bose = np.asarray(n_bose, dtype=np.int64)
        self.n_fermi = np.asarray(n_fermi, dtype=np.int64)
        uf = basis.u(self.x_nodes, lmax=L) * np.sqrt(2.0 / beta)                 # [nts, L]
        ub = basis.uhat(2 * self.n_bose, lmax=L) * np.sqrt(beta / 2.0)           # [nw_b, L]
        uff = basis.uhat(2 * self.n_fermi + 1, lmax=L) * np.sqrt(beta / 2.0)     # [nw, L]
        pf = np.linalg.pinv(uf, rcond=1e-13)
        self.T_bt = ub @ pf
        self.T_tb = uf @ np.linalg.pinv(ub, rcond=1e-13)
        self.T_ft = uff @ pf
        self.T_tf = uf @ np.linalg.pinv(uff, rcond=1e-13)
        self.cond_tau = np.linalg.cond(uf)
        self.cond_bose = np.linalg.cond(ub)
        self.cond_fermi = np.linalg.cond(uff)

    @property
    def nts(self) -> int:
        return len(self.tau)

    @property
    def nw_b(self) -> int:
        return len(self.n_bose)

    @property
    def nw(self) -> int:
        return len(self.n_fermi)

    @classmethod
    def from_grid(cls, data: dict, beta: float) -> "Transformer":
        b = Basis(lam=float(data["lam"]), edges=data["edges"], coef=data["coef"], s=data["s"])
        return cls(b, int(data["L"]), beta, data["x_nodes"], data["n_bose"], data["n_fermi"])

    @classmethod
    def from_nodes(cls, data: dict, beta: float, tau_nodes: np.ndarray, L: int | None = None) -> "Transformer":
        """Same basis and Matsubara sets, but fermionic tau nodes supplied by the caller."""
        b = Basis(lam=float(data["lam"]), edges=data["edges"], coef=data["coef"], s=data["s"])
        x = 2.0 * np.asarray(tau_nodes, dtype=np.float64) / beta - 1.0
        return cls(b, int(data["L"]) if L is None else L, beta, x, data["n_bose"], data["n_fermi"])

    # --- transforms on arrays whose leading axis is tau / frequency ---
    def tau_f_to_w_b(self, f_t: np.ndarray) -> np.ndarray:
        return np.tensordot(self.T_bt, f_t, axes=(1, 0))

    def w_b_to_tau_f(self, f_w: np.ndarray) -> np.ndarray:
        return np.tensordot(self.T_tb, f_w, axes=(1, 0))

    def tau_to_omega(self, f_t: np.ndarray) -> np.ndarray:
        return np.tensordot(self.T_ft, f_t, axes=(1, 0))

    def omega_to_tau(self, f_w: np.ndarray) -> np.ndarray:
        return np.tensordot(self.T_tf, f_w, axes=(1, 0))
