"""Pins for the oracle's scalar energy terms: barrier, friction mollifier,
friction gradient/Hessian, orthogonality energy.

Each check is against an independent statement: the alternative barrier
derivative formulas of test/google_test/test.cpp:148-166 (algebraically equal
to model/barrier.cpp:16-28), central finite differences (the reference's own
tolerance is 1e-3, iAABB/test.cpp:497-512; ours are tighter), and closed-form
properties of the C1 friction mollifier (IPC, Li et al. 2020).
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from adb_amd.scenes import PCG32

DH, KAP = 1e-3, 1e-4


def barrier_closed(d, dh, k):
    # B(d) = -k (d - dh)^2 ln(d/dh) / dh^2 for d < dh
    if d >= dh:
        return 0.0
    return -k * (d - dh) ** 2 * math.log(d / dh) / dh ** 2


def barrier_grad_gtest(d, dh, k):
    # test/google_test/test.cpp:148-156
    if d >= dh:
        return 0.0
    t2 = d - dh
    return k * (t2 * math.log(d / dh) * -2.0 - (t2 * t2) / d) / (dh * dh)


def barrier_hess_gtest(d, dh, k):
    # test/google_test/test.cpp:158-166
    if d >= dh:
        return 0.0
    t2 = d - dh
    return k * ((math.log(d / dh) * -2.0 - t2 * 4.0 / d) + 1.0 / (d * d) * (t2 * t2)) / (dh * dh)


@pytest.mark.parametrize("d", [1e-9, 1e-7, 1e-5, 3e-4, 9.99e-4, 1.2e-3])
def test_barrier_closed_forms(d):
    B, B1, B2 = O.barrier(d, DH, KAP)
    assert B == pytest.approx(barrier_closed(d, DH, KAP), rel=1e-13, abs=0)
    assert B1 == pytest.approx(barrier_grad_gtest(d, DH, KAP), rel=1e-12, abs=1e-300)
    assert B2 == pytest.approx(barrier_hess_gtest(d, DH, KAP), rel=1e-12, abs=1e-300)


@pytest.mark.parametrize("d", [1e-6, 1e-5, 1e-4, 5e-4, 9e-4])
def test_barrier_finite_difference(d):
    h = d * 1e-6
    Bp = O.barrier(d + h, DH, KAP)
    Bm = O.barrier(d - h, DH, KAP)
    B, B1, B2 = O.barrier(d, DH, KAP)
    assert B1 == pytest.approx((Bp[0] - Bm[0]) / (2 * h), rel=1e-6)
    assert B2 == pytest.approx((Bp[1] - Bm[1]) / (2 * h), rel=1e-6)


def test_barrier_vanishes_at_threshold():
    # C2 at d_hat: B, B' and B'' are all 0 there (the barrier is active only for d < d_hat)
    assert list(O.barrier(DH, DH, KAP)) == [0.0, 0.0, 0.0]
    B, B1, B2 = O.barrier(DH * (1 - 1e-9), DH, KAP)
    assert abs(B) < 1e-20 and abs(B1) < 1e-10 and B2 > 0


def test_friction_mollifier_closed_forms():
    eps = 1e-5
    for x in [0.0, 1e-7, 5e-6, 9.999e-6, 1e-5, 2e-5, 1.0]:
        f0 = O.f0(x, eps)
        f1x = O.f1_over_x(x, eps)
        if x >= eps:
            assert f0 == x
            assert f1x == pytest.approx(1.0 / x, rel=1e-15)
        else:
            # f0(x) = -x^3/(3 eps^2) + x^2/eps + eps/3;  f1 = f0' ; f1/x = (2 - x/eps)/eps
            assert f0 == pytest.approx(-x ** 3 / (3 * eps ** 2) + x ** 2 / eps + eps / 3, rel=1e-14)
            assert f1x == pytest.approx((2 - x / eps) / eps, rel=1e-14)
    # C1 continuity at eps
    assert O.f0(eps * (1 - 1e-12), eps) == pytest.approx(eps, rel=1e-10)
    h = 1e-12
    assert (O.f0(eps + h, eps) - O.f0(eps - h, eps)) / (2 * h) == pytest.approx(1.0, rel=1e-4)


def _friction_energy(x, T, lam, mu, eps):
    u = T.T @ x
    return mu * lam * O.f0(float(np.hypot(u[0], u[1])), eps)


@pytest.mark.parametrize("scale", [3.0, 0.3, 0.05])   # u >= eps, u < eps branches
def test_friction_gradient_hessian_fd(scale):
    rng = PCG32(11)
    n, mu, lam, eps = 12, 0.5, 2.3, 1e-5
    for trial in range(5):
        T = np.array([[rng.uniform(-1, 1) for _ in range(2)] for _ in range(n)])
        x = np.array([rng.uniform(-1, 1) for _ in range(n)])
        u = T.T @ x
        x = x * (scale * eps / np.linalg.norm(u))
        u = T.T @ x
        g, H = O.friction(u, lam, T, mu, eps)
        # gradient of mu lam f0(|T^T x|) w.r.t. x
        h = 1e-4 * eps
        gfd = np.array([(_friction_energy(x + h * e, T, lam, mu, eps) - _friction_energy(x - h * e, T, lam, mu, eps)) / (2 * h)
                        for e in np.eye(n)])
        np.testing.assert_allclose(g, gfd, rtol=1e-5, atol=1e-9 * np.abs(g).max())
        # Hessian by FD of the gradient
        Hfd = np.zeros((n, n))
        for k in range(n):
            e = np.eye(n)[k]
            gp, _ = O.friction(T.T @ (x + h * e), lam, T, mu, eps)
            gm, _ = O.friction(T.T @ (x - h * e), lam, T, mu, eps)
            Hfd[:, k] = (gp - gm) / (2 * h)
        np.testing.assert_allclose(H, Hfd, rtol=1e-4, atol=1e-6 * np.abs(H).max())
        # the smooth-friction Hessian is PSD in both branches
        assert np.linalg.eigvalsh(H).min() > -1e-9 * np.abs(H).max()


def test_friction_zero_velocity_branch():
    # u == 0 <= max_uk: D = mu lam f1 T T^T with f1/x(0) = 2/eps (model/friction.cpp:63-65)
    T = np.arange(24, dtype=float).reshape(12, 2) / 10
    g, H = O.friction([0.0, 0.0], 1.7, T, 0.4, 1e-5, max_uk=0.0)
    assert np.all(g == 0)
    np.testing.assert_allclose(H, 0.4 * 1.7 * (2 / 1e-5) * T @ T.T, rtol=1e-14)


def _ortho_energy(q, kap):
    A = np.asarray(q).reshape(4, 3)[1:].T   # columns a1, a2, a3
    return kap * np.sum((A.T @ A - np.eye(3)) ** 2)


def test_ortho_energy_gradient_hessian():
    rng = PCG32(5)
    kap = 1e9
    for trial in range(4):
        q = np.array([rng.uniform(-1, 1) for _ in range(12)])
        q[3:] = np.eye(3).T.reshape(-1) + 0.05 * q[3:]
        E, g, H = O.ortho(q, kap)
        assert E == pytest.approx(_ortho_energy(q, kap), rel=1e-12)
        h = 1e-6
        gfd = np.array([(_ortho_energy(q + h * e, kap) - _ortho_energy(q - h * e, kap)) / (2 * h) for e in np.eye(12)])
        np.testing.assert_allclose(g, gfd, rtol=1e-6, atol=1e-6 * np.abs(g).max())
        Hfd = np.zeros((12, 12))
        for k in range(12):
            e = np.eye(12)[k]
            Hfd[:, k] = (O.ortho(q + h * e, kap)[1] - O.ortho(q - h * e, kap)[1]) / (2 * h)
        np.testing.assert_allclose(H, Hfd, rtol=1e-6, atol=1e-6 * np.abs(H).max())
        np.testing.assert_allclose(H, H.T, rtol=0, atol=0)
    # a rotation is a minimum: zero energy and gradient
    R = np.linalg.qr(np.array([[1.0, 2, 3], [0.5, -1, 2], [2, 1, -1]]))[0]
    q = np.concatenate([[0.3, 0.2, 0.1], R[:, 0], R[:, 1], R[:, 2]])
    E, g, H = O.ortho(q, kap)
    assert E < 1e-12 * kap and np.abs(g).max() < 1e-6 * kap
