"""Two GW self-consistency iterations with every step on the device -- needs an MI355X.

G -> Sigma1 (gwhip_hf, from dm = -2 G(beta)) and Sigma(tau) (gwhip_eval_sigma, the GW kernel) -> mu
(gwhip_dyson_find_mu) -> G (gwhip_dyson_compute_G), the loop that green-sc runs around hf_solver::solve,
gw_solver::solve and dyson::solve (src/hf_solver.cpp:9-16, src/gw_solver.cpp:30-32, src/dyson.cpp:101-236), on
a small seeded system (nk = 2, nao = 4, naux = 10, ns = 1, beta = 20, IR Lambda = 1e3, S = I, half
filling).  Each device step is checked against its oracle on the device's own input: Sigma1 against
oracle.hf_oracle.hf_fock and Sigma(tau) against oracle.gw_oracle.gw_sigma within 1e-10, the chemical
potential against oracle.dyson.find_mu on the same self-energies within 1e-8 (both bisect to a relative
electron-count tolerance of 1e-9), G against oracle.dyson.compute_G at the device's mu within 1e-10.
The first G is the Dyson G of H alone (Sigma1 = Sigma(tau) = 0) at its own mu.  The same loop with
the GW step in split precision keeps Sigma within 1e-8 of the oracle's FP64 Sigma of the same G.
"""
import numpy as np
import pytest

from gwhip import inputs
from gwhip import lib as gl
from helpers import rel_err, transformer
from oracle import dyson as od
from oracle import gw_oracle as go
from oracle import hf_oracle as ho

pytestmark = pytest.mark.gpu

NK, NAO, NQ, NS, BETA, LAM = 2, 4, 10, 1, 20.0, 1e3
NEL = 4.0  # two of the four bands of each k filled, both spins


def _system(T):
    rng = np.random.default_rng(41)
    c = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    H = c(NS, NK, NAO, NAO) * 0.3
    H = 0.5 * (H + np.conj(np.swapaxes(H, -1, -2)))
    S = np.broadcast_to(np.eye(NAO), (NS, NK, NAO, NAO)).astype(np.complex128).copy()
    V = inputs.generate_V(41, NK, NQ, NAO, 0.1)
    return H, S, V


@pytest.mark.parametrize("precision", [gl.PREC_FP64, gl.PREC_SPLIT])
def test_two_sc_iterations_on_device(precision):
    T = transformer(LAM, BETA)
    wn = (2 * T.n_fermi + 1) * np.pi / BETA
    H, S, V = _system(T)
    zero1 = np.zeros((NS, NK, NAO, NAO), np.complex128)
    k = gl.GwHipKernel(NK, NAO, NAO, NS, NQ, T.nts, T.nw_b, BETA, mesh=(2, 1, 1))
    d = gl.Dyson(T.nts, len(wn), NS, NK, NK, NAO, NAO)
    try:
        k.set_grid(T.T_bt, T.T_tb)
        k.upload_V(V)
        d.set_grid(T.T_ft, T.T_tf, wn)
        d.set_system(H, S)

        def step_mu_G(Sigma1, Sigma_tau, mu0):
            nel, mu = d.find_mu(Sigma1, Sigma_tau, mu0, NEL, 1e-9)
            eig = od.eigenspectra(H, S, Sigma1, Sigma_tau, T.T_ft)
            N = lambda m: od.number_of_electrons(m, eig, wn, T.T_tf, [1.0] * NK, NK, NS)
            _, mu_or = od.find_mu(mu0, NEL, 1e-9, N)
            G = d.compute_G(Sigma1, Sigma_tau, mu)
            G_or = od.compute_G(H, S, Sigma1, Sigma_tau, mu, T.T_ft, T.T_tf, wn)
            print(f"mu {mu:.12f} (oracle {mu_or:.12f}), N {nel:.12f}; G {rel_err(G, G_or):.2e}", flush=True)
            assert abs(mu - mu_or) < 1e-8 and abs(nel - NEL) < 1e-8
            assert rel_err(G, G_or) < 1e-10
            return G, mu

        G, mu = step_mu_G(zero1, np.zeros((T.nts, NS, NK, NAO, NAO), np.complex128), 0.0)
        for it in range(2):
            dm = ho.dm_from_G(G[-1], NS, False)
            Sigma1 = k.hf(dm, S, 0.0)
            Sigma1_or = ho.hf_fock(dm, V, S, 0.0, NK, NAO, NAO, NS, NQ)
            print(f"iteration {it}: Sigma1 {rel_err(Sigma1, Sigma1_or):.2e} vs oracle", flush=True)
            assert rel_err(Sigma1, Sigma1_or) < 1e-10
            Sigma = k.solve(G, precision)
            Sigma_or = go.gw_sigma(G, V, T.T_bt, T.T_tb, NK, NAO, NAO, NS, NQ, mesh=(2, 1, 1))
            print(f"iteration {it}: Sigma {rel_err(Sigma, Sigma_or):.2e} vs oracle", flush=True)
            assert rel_err(Sigma, Sigma_or) < (1e-10 if precision == gl.PREC_FP64 else 1e-8)
            G_new, mu = step_mu_G(Sigma1, Sigma, mu)
            assert rel_err(G_new, G) > 1e-6  # the self-energy moved G
            G = G_new
    finally:
        d.close()
        k.close()
