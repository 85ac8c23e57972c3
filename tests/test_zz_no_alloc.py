"""gwhip_eval_sigma_device_noalloc against gwhip_eval_sigma_device -- needs an MI355X.

Two calls in a row through the C ABI share the library's one device workspace: first a system whose
tau chunks do not all fit it (nk = 27, nao = 26, naux = 150, IR Lambda = 1e4: the workspace holds
two or three tau points per chunk and one q per Dyson chunk, the handle's own workspace all 50 and
every q), then a smaller one (nk = 8, nao = 8, naux = 40).  Each must be bit-identical to the allocating entry point on the same handle, in FP64 and
in split precision (chunking only changes which launch a tau point or q runs in, never its
arithmetic).  A system that does not fit the workspace at one tau point is refused with
GWHIP_ERR_UNSUPPORTED, and the allocating entry point still works after the no-allocation calls.
"""
import numpy as np
import pytest

from gwhip import inputs, irgrid
from gwhip import lib as gl

pytestmark = pytest.mark.gpu


def _system(nk, nao, NQ, lam, beta, seed):
    T = irgrid.Transformer.from_grid(irgrid.load_grid(lam), beta)
    n = round(nk ** (1 / 3))
    k = gl.GwHipKernel(nk, nao, nao, 1, NQ, T.nts, T.nw_b, beta, mesh=(n, n, n))
    k.set_grid(T.T_bt, T.T_tb)
    k.generate_V(seed, 0.1)
    G = inputs.generate_G(seed, nk, nao, 1, T.tau, beta)
    return k, G


def _both(k, G, precision):
    import torch

    dev = torch.device("cuda:0")
    G_d = torch.from_numpy(G.view(np.float64)).to(dev)
    S_ref = torch.empty_like(G_d)
    S_na = torch.full_like(G_d, float("nan"))  # the entry point must write all of Sigma
    torch.cuda.synchronize()
    k.solve_device(G_d.data_ptr(), S_ref.data_ptr(), precision)  # builds the split slice stores
    k.solve_device_noalloc(G_d.data_ptr(), S_na.data_ptr(), precision)
    torch.cuda.synchronize()
    return S_ref.cpu().numpy(), S_na.cpu().numpy(), G_d


@pytest.mark.parametrize("precision", [gl.PREC_FP64, gl.PREC_SPLIT])
def test_noalloc_matches_allocating_entry(precision):
    import torch

    for nk, nao, NQ, lam, beta in [(27, 26, 150, 1e4, 100.0), (8, 8, 40, 1e3, 50.0)]:
        print(f"noalloc: precision {precision}, nk {nk} nao {nao} naux {NQ}", flush=True)
        k, G = _system(nk, nao, NQ, lam, beta, seed=nk)
        try:
            ref, got, G_d = _both(k, G, precision)
            assert np.isfinite(got).all()
            assert np.array_equal(ref, got), f"max diff {np.abs(ref - got).max():.3e}"
            # the handle's own workspace is back in place after the no-allocation call
            S2 = torch.empty_like(G_d)
            k.solve_device(G_d.data_ptr(), S2.data_ptr(), precision)
            assert np.array_equal(ref, S2.cpu().numpy())
        finally:
            k.close()


def test_noalloc_refuses_what_does_not_fit():
    """nk = 64, nao = 26, naux = 300 needs about 0.64 GB at one tau point and one q."""
    import torch

    T = irgrid.Transformer.from_grid(irgrid.load_grid(1e4), 100.0)
    k = gl.GwHipKernel(64, 26, 26, 1, 300, T.nts, T.nw_b, 100.0, mesh=(4, 4, 4))
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.generate_V(0, 0.1)
        G = torch.zeros(T.nts * 64 * 26 * 26 * 2, dtype=torch.float64, device="cuda:0")
        S = torch.empty_like(G)
        with pytest.raises(gl.GwhipError) as e:
            k.solve_device_noalloc(G.data_ptr(), S.data_ptr(), gl.PREC_FP64)
        assert e.value.status == gl.ERR_UNSUPPORTED
    finally:
        k.close()
