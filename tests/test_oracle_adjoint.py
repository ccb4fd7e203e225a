"""The adjoint of adb_amd/adjoint.py checked on the CPU with the oracle world
(the same Rollout driver, the oracle's exact Hessian and sparse Cholesky):
gradient of a tail-link centroid loss with respect to the initial velocities
and positions of a 3-link torus chain against central finite differences."""
import numpy as np
import pytest
import torch

import oracle_lib as O
from adb_amd import adjoint, scenes
from adb_amd.capi import default_params


def chain_params():
    # tight Newton tolerance: finite differences of the rollout resolve the
    # gradient only as far as each step is converged
    return default_params(d_hat=1e-5, kappa=1e-4, mu=0.0, dt=0.01, enable_ground=0, tol=1e-12)


def make(n_links, steps, every):
    sc = scenes.torus_chain(n_links, seed=4)
    # a sideways kick so links collide within the short horizon
    sc["dqdt"][1:, 0] = 0.3 * np.sin(np.arange(1, n_links))
    p = chain_params()
    sim = adjoint.Rollout(sc, p, steps, every, world=O.OracleWorld(sc, p))
    return sc, sim


def loss_of(sc, q, target):
    c = adjoint.centroid(sc, q, len(sc["mass"]) - 1)
    return 0.5 * ((c - target) ** 2).sum()


def test_adjoint_matches_finite_differences():
    sc, sim = make(3, 12, 5)
    q0 = torch.tensor(sc["q0"], dtype=torch.float64, requires_grad=True)
    v0 = torch.tensor(sc["dqdt"], dtype=torch.float64, requires_grad=True)
    target = adjoint.centroid(sc, q0.detach(), 2) + torch.tensor([0.05, -0.05, 0.02], dtype=torch.float64)
    q, qd = adjoint.rollout(q0, v0, sim)
    L = loss_of(sc, q, target) + 0.01 * (qd[2, :3] ** 2).sum()
    L.backward()
    assert max(sim.stats) >= 1
    rng = np.random.default_rng(1)

    def L_at(qq, vv):
        q, qd, _ = sim.forward(qq, vv)
        q, qd = torch.from_numpy(q), torch.from_numpy(qd)
        return float(loss_of(sc, q, target) + 0.01 * (qd[2, :3] ** 2).sum())

    for which in ("v", "q"):
        d = rng.normal(size=(3, 12))
        d[0] = 0.0                                      # the pinned link
        if which == "q":
            d[:, 3:] *= 0.1
        eps = 1e-6
        base_q, base_v = sc["q0"].copy(), sc["dqdt"].copy()
        if which == "v":
            fd = (L_at(base_q, base_v + eps * d) - L_at(base_q, base_v - eps * d)) / (2 * eps)
            ad = float((v0.grad.numpy() * d).sum())
        else:
            fd = (L_at(base_q + eps * d, base_v) - L_at(base_q - eps * d, base_v)) / (2 * eps)
            ad = float((q0.grad.numpy() * d).sum())
        assert ad == pytest.approx(fd, rel=1e-4), (which, ad, fd)
