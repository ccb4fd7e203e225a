"""The differentiable step on the GPU (adb_amd/adjoint.py over
adb_adjoint_solve): gradients of a tail-link centroid loss through a torus
chain's rollout.

* 3-link chain, 12 steps: the GPU adjoint gradient (w.r.t. initial velocities
  and positions) equals the oracle adjoint's (exact Hessian, sparse Cholesky)
  within 1e-6 relative, and central finite differences within 1e-4.
* The 64-link chain (24 x 12 tori, R 0.1, r 0.025, alternate links turned 90
  degrees, top link pinned, no ground), 100 steps with a checkpoint every 10:
  the gradient w.r.t. the 64 x 12 initial velocities against central finite
  differences along two random directions, within 1e-4 relative.
"""
import numpy as np
import pytest

import oracle_lib as O
from adb_amd import adjoint, scenes
from adb_amd.capi import default_params

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def chain_params():
    # sqrt(d_hat) = 3.2 mm, below the 25 mm tube radius; frictionless.  The
    # Newton tolerance is tight because finite differences of the rollout
    # resolve the gradient only as far as each step is converged (1e-10 leaves
    # ~1e-4 relative noise in the smaller position derivatives)
    return default_params(d_hat=1e-5, kappa=1e-4, mu=0.0, dt=0.01, enable_ground=0, tol=1e-12)


def chain(n):
    sc = scenes.torus_chain(n, seed=4)
    sc["dqdt"][1:, 0] = 0.3 * np.sin(np.arange(1, n))
    return sc


def grads(sim, sc, target, tail):
    q0 = torch.tensor(sc["q0"], dtype=torch.float64, requires_grad=True)
    v0 = torch.tensor(sc["dqdt"], dtype=torch.float64, requires_grad=True)
    q, qd = adjoint.rollout(q0, v0, sim)
    L = 0.5 * ((adjoint.centroid(sc, q, tail) - target) ** 2).sum()
    L.backward()
    return float(L), q0.grad.numpy(), v0.grad.numpy()


def loss_at(sim, sc, target, tail, q0, v0):
    q, _, _ = sim.forward(q0, v0)
    c = adjoint.centroid(sc, torch.from_numpy(q), tail)
    return float(0.5 * ((c - target) ** 2).sum())


def test_adjoint_small_chain_vs_oracle_and_fd():
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    sc = chain(3)
    p = chain_params()
    target = adjoint.centroid(sc, torch.tensor(sc["q0"]), 2) + torch.tensor([0.05, -0.05, 0.02], dtype=torch.float64)
    gsim = adjoint.Rollout(sc, p, 12, 5)
    osim = adjoint.Rollout(sc, p, 12, 5, world=O.OracleWorld(sc, p))
    Lg, gq, gv = grads(gsim, sc, target, 2)
    Lo, oq, ov = grads(osim, sc, target, 2)
    assert Lg == pytest.approx(Lo, rel=1e-10)
    np.testing.assert_allclose(gv, ov, rtol=0, atol=1e-6 * np.abs(ov).max())
    np.testing.assert_allclose(gq, oq, rtol=0, atol=1e-6 * np.abs(oq).max())
    d = np.random.default_rng(3).normal(size=(3, 12))
    d[0] = 0.0
    eps = 1e-6
    fd = (loss_at(gsim, sc, target, 2, sc["q0"], sc["dqdt"] + eps * d)
          - loss_at(gsim, sc, target, 2, sc["q0"], sc["dqdt"] - eps * d)) / (2 * eps)
    assert float((gv * d).sum()) == pytest.approx(fd, rel=1e-4)


def test_adjoint_full_chain_fd():
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    sc = scenes.torus_chain(64, seed=4)
    p = chain_params()
    sim = adjoint.Rollout(sc, p, 100, 10)
    tail = 63
    target = adjoint.centroid(sc, torch.tensor(sc["q0"]), tail) + torch.tensor([0.05, -0.05, 0.02], dtype=torch.float64)
    L, gq, gv = grads(sim, sc, target, tail)
    assert np.abs(gv[1:]).max() > 0 and np.all(gv[0] == 0)
    rng = np.random.default_rng(64)
    for _ in range(2):
        d = rng.normal(size=(64, 12))
        d[0] = 0.0
        d[:, 3:] *= 0.1
        # the 100-step rollout is stiff (contacts come and go): the central
        # difference's truncation error is 3e-2, 7e-3, 3e-4 of the derivative
        # at eps 1e-4, 1e-5, 1e-6 in one direction, 1e-6 at eps 1e-7, where
        # rounding (Newton tolerance 1e-12) is still far below it
        eps = 1e-7
        fd = (loss_at(sim, sc, target, tail, sc["q0"], sc["dqdt"] + eps * d)
              - loss_at(sim, sc, target, tail, sc["q0"], sc["dqdt"] - eps * d)) / (2 * eps)
        assert float((gv * d).sum()) == pytest.approx(fd, rel=1e-4)
