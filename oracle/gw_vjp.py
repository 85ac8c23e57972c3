"""Reference gradients of Sigma[G] -- TEST INFRASTRUCTURE ONLY.

The reference has no derivative of the GW self-energy; this module supplies the two checkers the
device VJP is held to, and nothing in the library imports it:

* ``sigma_dense`` -- the forward chain of ``gw_oracle.gw_sigma`` (``src/gw_cpu_kernel.cpp:31-373``)
  written as dense complex128 torch contractions on the CPU, so that ``torch.autograd`` yields
  the exact vector-Jacobian product (``vjp_autograd``).  It is pinned to ``gw_oracle.gw_sigma``
  on the same inputs.
* ``vjp_manual`` -- the hand-derived backward pass, step for step the algorithm the HIP VJP
  runs (per q: Sigma part -> Dyson adjoint -> P0 part), in numpy.  It is pinned to
  ``vjp_autograd`` and to central finite differences of ``gw_oracle.gw_sigma``.

Convention (the same as torch's): for a cotangent Sbar, the VJP is the Gbar with
Re<Sbar, dSigma> = Re<Gbar, dG> for every dG, <a, b> = sum conj(a) b.
"""
from __future__ import annotations

import numpy as np

from oracle import gw_oracle as go


def v_full(Vp: np.ndarray, nk: int) -> np.ndarray:
    """V(k1, k2) for every ordered pair: [nk, nk, NQ, nao, nao]."""
    return np.stack([np.stack([go.v_pair(Vp, k1, k2) for k2 in range(nk)]) for k1 in range(nk)])


def _blocks(nao, nso, ns):
    X2C = nso != nao
    pns = nso // nao
    pseudo_ns = 4 if X2C else ns
    pref = 1.0 if (ns == 2 or X2C) else 2.0
    out = []
    for s in range(pseudo_ns):
        out.append((s // (pns * pns), (s // pns) % pns, s % pns))
    return out, pref


def _herm(x, xp):
    return 0.5 * (x + xp.conj(x.swapaxes(-1, -2)))


def sigma_dense(G, Vf, T_bt, T_tb, nk, nao, nso, ns, NQ, mesh):
    """Dense torch forward (complex128, autograd-able).  G: torch [nts, ns, nk, nso, nso]."""
    import torch

    nts = G.shape[0]
    half = nts // 2
    blocks, pref = _blocks(nao, nso, ns)
    Vf = torch.as_tensor(Vf)
    Tbt = torch.as_tensor(T_bt)
    Ttb = torch.as_tensor(T_tb)
    I = torch.eye(NQ, dtype=torch.complex128)
    kp = [[go.k_add(k, q, mesh, +1) for k in range(nk)] for q in range(nk)]
    km = [[go.k_add(k, q, mesh, -1) for k in range(nk)] for q in range(nk)]
    Sb = {blk: 0 for blk in blocks}
    tt = list(range(nts - 1, nts - 1 - half, -1))
    for q in range(nk):
        v = Vf[torch.arange(nk), torch.tensor(kp[q])]                       # [k, Q, p, m] = V(k, k+q)
        P0 = torch.zeros((half, NQ, NQ), dtype=torch.complex128)
        for (is_, a, b) in blocks:
            Gb = G[tt, is_][:, :, b * nao:(b + 1) * nao, a * nao:(a + 1) * nao]          # [t, k, x, p]
            Gq = G[:half, is_][:, kp[q]][:, :, a * nao:(a + 1) * nao, b * nao:(b + 1) * nao]  # [t, k, m, n]
            Z = torch.einsum("tkxp,kQpm->tkQxm", Gb, v)
            Z = torch.einsum("tkQxm,tkmn->tkQxn", Z, Gq)
            P0 = P0 + torch.einsum("tkQxn,kRxn->tQR", Z, v.conj())
        P0 = _herm(P0 * (-pref / nk), torch)
        P0full = torch.cat([P0, torch.flip(P0, dims=[0])], dim=0)
        Pw = torch.einsum("wt,tQR->wQR", Tbt, P0full)
        Binv = torch.linalg.inv(I - Pw)
        Pw_t = _herm(Binv, torch) @ Pw
        Pt = torch.einsum("tw,wQR->tQR", Ttb, Pw_t)
        vm = Vf[torch.arange(nk), torch.tensor(km[q])]                      # [k, Q, i, m] = V(k, k-q)
        for (is_, a, b) in blocks:
            Gm = G[:, is_][:, km[q]][:, :, a * nao:(a + 1) * nao, b * nao:(b + 1) * nao]   # [t, k, m, n]
            T = torch.einsum("kQim,tkmn->tkQin", vm, Gm)
            Y = torch.einsum("tkQin,tQR->tkRin", T, Pt)
            Sb[(is_, a, b)] = Sb[(is_, a, b)] + torch.einsum("tkRin,kRjn->tkij", Y, vm.conj()) * (-1.0 / nk)
    pns = nso // nao
    spins = []
    for is_ in range(ns):
        rows = []
        for a in range(pns):
            rows.append(torch.cat([Sb[(is_, a, b)] for b in range(pns)], dim=-1))
        spins.append(torch.cat(rows, dim=-2))
    Sigma = torch.stack(spins, dim=1)
    return Sigma


def vjp_autograd(G, Sbar, Vp, T_bt, T_tb, nk, nao, nso, ns, NQ, mesh):
    """Exact VJP of ``sigma_dense`` by torch autograd (CPU, complex128)."""
    import torch

    Gt = torch.tensor(G, dtype=torch.complex128, requires_grad=True)
    S = sigma_dense(Gt, v_full(Vp, nk), T_bt, T_tb, nk, nao, nso, ns, NQ, mesh)
    (g,) = torch.autograd.grad(S, Gt, grad_outputs=torch.tensor(Sbar, dtype=torch.complex128))
    return S.detach().numpy(), g.numpy()


def vjp_manual(G, Sbar, Vp, T_bt, T_tb, nk, nao, nso, ns, NQ, mesh, return_parts=False, plant=None):
    """Hand-derived VJP of Sigma[G] in the order the device runs it.

    ``plant`` (tests only) leaves out one term: "no_inverse_adjoint" (the -Binv^H (.) Binv^H
    part of the Dyson adjoint) or "no_gbeta_term" (the G(beta - tau) part of the P0 adjoint).

    Per q, with c = -1/nk, alpha = -pref/nk, v = V(k, k-q) (Sigma part) or V(k, k+q) (P0 part):
      Sigma part (every tau, spin block s):
        W_Q'  = Sbar_s v_Q'                      Y_Q = sum_Q' conj(Pt[Q,Q']) W_Q'
        Gbar(tau, k-q)_s += c sum_Q v_Q^H Y_Q    T_Q = v_Q G(tau, k-q)_s
        Ptbar[Q,Q'] += c sum_k sum_(i,n) conj(T_Q[i,n]) W_Q'[i,n]
      Dyson adjoint (every Omega):
        Pwbar_t = sum_tau conj(T_tb[tau, Omega]) Ptbar(tau)
        P0wbar = Ainv^H Pwbar_t + Binv^H H(Pwbar_t P0w^H) Binv^H,  Ainv = H(Binv), Binv = (I - P0w)^-1
        P0hbar(t) = sum_Omega conj(Tfold[Omega, t]) P0wbar(Omega),  Tfold = T_bt[:, t] + T_bt[:, nts-1-t]
        Pbar(t) = H(P0hbar(t))
      P0 part (half grid t, every k, spin block s):
        Zbar_Q = alpha sum_Q' Pbar[Q,Q'] v_Q'
        Gbar(beta - tau, k)_(b,a) += sum_Q Zbar_Q Gq^H v_Q^H
        Gbar(tau, k + q)_(a,b)   += sum_Q v_Q^H Gb^H Zbar_Q
    """
    G = np.asarray(G, np.complex128)
    Sbar = np.asarray(Sbar, np.complex128)
    nts = G.shape[0]
    half = nts // 2
    blocks, pref = _blocks(nao, nso, ns)
    Vf = v_full(Vp, nk)
    Gbar = np.zeros_like(G)
    c = -1.0 / nk
    alpha = -pref / nk
    Tfold = T_bt[:, :half] + T_bt[:, ::-1][:, :half]
    I = np.eye(NQ)
    parts = []
    for q in range(nk):
        kp = [go.k_add(k, q, mesh, +1) for k in range(nk)]
        km = [go.k_add(k, q, mesh, -1) for k in range(nk)]
        # forward quantities of this q (recomputed)
        P0 = np.zeros((nts, NQ, NQ), np.complex128)
        go.eval_P0_tilde(P0, G, Vp, q, nk, nao, nso, ns, NQ, nso != nao, mesh)
        go.symmetrize_P0(P0, nk)
        P0w = np.tensordot(T_bt, P0, axes=(1, 0))
        Binv = np.linalg.inv(I[None] - P0w)
        Ainv = _herm(Binv, np)
        Pw_t = Ainv @ P0w
        Pt = np.tensordot(T_tb, Pw_t, axes=(1, 0))
        # Sigma part
        Ptbar = np.zeros((nts, NQ, NQ), np.complex128)
        for k in range(nk):
            v = Vf[k, km[k]]                                            # [Q, i, m]
            for (is_, a, b) in blocks:
                sa, sb = slice(a * nao, (a + 1) * nao), slice(b * nao, (b + 1) * nao)
                Sb = Sbar[:, is_, k][:, sa, sb]                         # [t, i, j]
                Gm = G[:, is_, km[k]][:, sa, sb]                        # [t, m, n]
                W = np.einsum("tij,Rjn->tRin", Sb, v)
                Y = np.einsum("tQR,tRin->tQin", np.conj(Pt), W)
                Gbar[:, is_, km[k], sa, sb] += c * np.einsum("Qim,tQin->tmn", np.conj(v), Y)
                T = np.einsum("Qim,tmn->tQin", v, Gm)
                Ptbar += c * np.einsum("tQin,tRin->tQR", np.conj(T), W)
        # Dyson adjoint
        Pwbar_t = np.tensordot(np.conj(T_tb).T, Ptbar, axes=(1, 0))
        Ainvbar = Pwbar_t @ np.conj(np.swapaxes(P0w, -1, -2))
        Binvh = np.conj(np.swapaxes(Binv, -1, -2))
        P0wbar = np.conj(np.swapaxes(Ainv, -1, -2)) @ Pwbar_t
        if plant != "no_inverse_adjoint":
            P0wbar = P0wbar + Binvh @ _herm(Ainvbar, np) @ Binvh
        P0hbar = np.tensordot(np.conj(Tfold).T, P0wbar, axes=(1, 0))   # [half, Q, Q']
        Pbar = _herm(P0hbar, np)
        # P0 part
        for k in range(nk):
            v = Vf[k, kp[k]]                                            # [Q, p, m]
            Zbar = alpha * np.einsum("tQR,Rxn->tQxn", Pbar, v)          # [t, Q, x, n]
            for (is_, a, b) in blocks:
                sa, sb = slice(a * nao, (a + 1) * nao), slice(b * nao, (b + 1) * nao)
                Gb = G[nts - 1 - np.arange(half), is_, k][:, sb, sa]    # [t, x, p]
                Gq = G[:half, is_, kp[k]][:, sa, sb]                    # [t, m, n]
                gb = np.einsum("tQxn,tmn,Qpm->txp", Zbar, np.conj(Gq), np.conj(v))
                gq = np.einsum("Qpm,txp,tQxn->tmn", np.conj(v), np.conj(Gb), Zbar)
                if plant != "no_gbeta_term":
                    Gbar[nts - 1 - np.arange(half), is_, k, sb, sa] += gb
                Gbar[:half, is_, kp[k], sa, sb] += gq
        if return_parts:
            parts.append(dict(Pt=Pt, Ptbar=Ptbar, Pbar=Pbar))
    if return_parts:
        return Gbar, parts
    return Gbar
