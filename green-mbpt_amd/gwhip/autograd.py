"""torch.autograd binding of the device Sigma[G] and its hand-written VJP (driver / test harness).

``sigma(G, kernel)`` returns Sigma(tau, k) for a complex128 CUDA tensor G of shape
[nts, ns, nk, nso, nso]; its backward calls ``gwhip_vjp_device``, never torch's own
differentiation of the contraction.  ``mode`` picks whether the backward recomputes
P0/P-tilde per q or reuses those the forward kept (``GwHipKernel.set_vjp_store(True)``).
"""
from __future__ import annotations

import torch

from gwhip import lib as gl


class _SigmaFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, G, kernel, precision, mode):
        if not (G.is_cuda and G.dtype == torch.complex128):
            raise TypeError("G must be a complex128 CUDA tensor")
        Gc = G.detach().contiguous()
        S = torch.empty_like(Gc)
        torch.cuda.current_stream(G.device).synchronize()  # the library runs on its own stream
        kernel.solve_device(Gc.data_ptr(), S.data_ptr(), precision)
        ctx.save_for_backward(Gc)
        ctx.kernel, ctx.mode = kernel, mode
        return S

    @staticmethod
    def backward(ctx, Sbar):
        (Gc,) = ctx.saved_tensors
        sb = Sbar.contiguous()
        Gbar = torch.empty_like(Gc)
        torch.cuda.current_stream(Gc.device).synchronize()
        ctx.kernel.vjp_device(Gc.data_ptr(), sb.data_ptr(), Gbar.data_ptr(), ctx.mode)
        return Gbar, None, None, None


def sigma(G: torch.Tensor, kernel: "gl.GwHipKernel", precision: int = gl.PREC_FP64,
          mode: int = gl.VJP_RECOMPUTE) -> torch.Tensor:
    return _SigmaFn.apply(G, kernel, precision, mode)
