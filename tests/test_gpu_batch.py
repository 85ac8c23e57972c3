"""Several uneven systems in one captured hipGraph (gwhip_batch_*) -- needs an MI355X.

* every system of a batch gets bit-for-bit the Sigma its own stand-alone kernel gets, in FP64 and
  in split precision, and FP64 matches the oracle to 1e-10;
* the graph is replayed several times with eager library and torch work and synchronisations in
  between, and every replay is checked against the eager result;
* new G / Sigma pointers or retuned launch parameters trigger a new capture, the same pointers
  do not; autotuning changes speed only.
"""
import numpy as np
import pytest

from gwhip import lib as gl
from helpers import rel_err, synthetic
from oracle import gw_oracle as go

pytestmark = pytest.mark.gpu

SYSTEMS = [  # uneven nk / nao / naux, ns = 2 with a spin-split Fock operator
    dict(nk=2, mesh=(2, 1, 1), nao=3, NQ=7),
    dict(nk=4, mesh=(2, 2, 1), nao=17, NQ=33),
    dict(nk=1, mesh=(1, 1, 1), nao=8, NQ=20),
]


def _make():
    out = []
    for i, c in enumerate(SYSTEMS):
        G, V, T, d = synthetic(c["nk"], c["nao"], c["NQ"], ns=2, beta=20.0, lam=1e3, mesh=c["mesh"], seed=30 + i,
                               sigma=0.3)
        out.append((G, V, T, d))
    return out


def _dims(d, T):
    return dict(nk=d["nk"], nao=d["nao"], nso=d["nso"], ns=d["ns"], NQ=d["NQ"], nts=T.nts, nw_b=T.nw_b,
                beta=d["beta"], mesh=d["mesh"])


def _standalone(G, V, T, d, precision):
    k = gl.GwHipKernel(d["nk"], d["nao"], d["nso"], d["ns"], d["NQ"], T.nts, T.nw_b, d["beta"], mesh=d["mesh"])
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.upload_V(V)
        return k.solve(G, precision)
    finally:
        k.close()


def _batch(systems):
    b = gl.Batch([_dims(d, T) for (_, _, T, d) in systems])
    for s, (G, V, T, d) in zip(b.systems, systems):
        s.set_grid(T.T_bt, T.T_tb)
        s.upload_V(V)
    return b


@pytest.mark.parametrize("precision", [gl.PREC_FP64, gl.PREC_SPLIT])
def test_batch_matches_standalone_and_replays(precision):
    import torch

    systems = _make()
    dev = torch.device("cuda:0")
    G_d = [torch.tensor(G, device=dev) for (G, _, _, _) in systems]
    S_d = [torch.empty_like(g) for g in G_d]
    ref = [_standalone(G, V, T, d, precision) for (G, V, T, d) in systems]
    b = _batch(systems)
    try:
        torch.cuda.synchronize()
        for rep in range(3):
            for s in S_d:
                s.fill_(float("nan"))
            torch.cuda.synchronize()
            b.eval_device([g.data_ptr() for g in G_d], [s.data_ptr() for s in S_d], precision)
            for i, (s, r) in enumerate(zip(S_d, ref)):
                got = s.cpu().numpy()
                assert np.array_equal(got, r), f"replay {rep}, system {i}: {rel_err(got, r):.2e}"
            # eager work between replays: a stand-alone evaluation and torch kernels
            _standalone(*systems[rep % len(systems)], precision)
            (G_d[0] * 2).sum().item()
        st = b.stats()
        assert st["captures"] == 1 and st["graph_nodes"] > 10
    finally:
        b.close()
    tol = 1e-10 if precision == gl.PREC_FP64 else 1e-8
    for (G, V, T, d), r in zip(systems, ref):
        S_o = go.gw_sigma(G, V, T.T_bt, T.T_tb, d["nk"], d["nao"], d["nso"], d["ns"], d["NQ"], mesh=d["mesh"])
        assert rel_err(r, S_o) < tol


def test_recapture_rules_and_autotune():
    import torch

    systems = _make()
    dev = torch.device("cuda:0")
    G_d = [torch.tensor(G, device=dev) for (G, _, _, _) in systems]
    S1 = [torch.empty_like(g) for g in G_d]
    S2 = [torch.empty_like(g) for g in G_d]
    b = _batch(systems)
    try:
        torch.cuda.synchronize()
        gp = [g.data_ptr() for g in G_d]
        b.eval_device(gp, [s.data_ptr() for s in S1])
        b.eval_device(gp, [s.data_ptr() for s in S1])
        assert b.stats()["captures"] == 1
        b.eval_device(gp, [s.data_ptr() for s in S2])
        assert b.stats()["captures"] == 2
        first = [s.cpu().numpy() for s in S1]
        assert all(np.array_equal(a, s.cpu().numpy()) for a, s in zip(first, S2))
        b.autotune(gp)
        for s in b.systems:
            t = s.tuning()
            assert t["p0_tile"] in (0, 1, 2) and t["sigma_tile"] in (0, 1, 2) and t["aux_per_wave"] in (1, 2, 4, 8)
            assert t["epi_form"] in (1, 2, 3)
        b.autotune(gp, gl.PREC_SPLIT)
        for s in b.systems:
            t = s.tuning()
            assert t["p0_group"] in (1, 2, 4, 8, 16, 32, 64) and t["sigma_group"] in (1, 2, 4, 8, 16, 32, 64)
        b.systems[1].set_tuning(p0_tile=1, sigma_tile=2, aux_per_wave=1, p0_group=4, sigma_group=8,
                                p0_split_tile=2, sigma_split_tile=2)
        b.eval_device(gp, [s.data_ptr() for s in S1])
        assert b.stats()["captures"] == 3
        for a, s in zip(first, S1):
            assert rel_err(s.cpu().numpy(), a) < 1e-14
    finally:
        b.close()


def test_batch_state_errors():
    import torch

    systems = _make()[:2]
    G_d = [torch.tensor(G, device="cuda:0") for (G, _, _, _) in systems]
    S_d = [torch.empty_like(g) for g in G_d]
    torch.cuda.synchronize()
    b = gl.Batch([_dims(d, T) for (_, _, T, d) in systems])
    try:
        b.systems[0].set_grid(systems[0][2].T_bt, systems[0][2].T_tb)
        b.systems[0].upload_V(systems[0][1])
        with pytest.raises(gl.GwhipError) as e:  # system 1 has no grid / V: refused before any launch
            b.eval_device([g.data_ptr() for g in G_d], [s.data_ptr() for s in S_d])
        assert e.value.status == gl.ERR_STATE
        with pytest.raises(gl.GwhipError) as e:
            b.eval_device([0, 0], [0, 0])
        assert e.value.status == gl.ERR_ARG
    finally:
        b.close()
