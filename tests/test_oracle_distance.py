"""Pins for the oracle's distance functions (the ipc-toolkit v1.2.1 calls at
model/geometry.cpp:101-112, model/abd.cpp:353-354, model/core_ipc.cpp:30-31,62-67;
ipc-toolkit is not vendored by the reference and cannot be fetched here).

* Typed squared distances agree with textbook closest-point routines
  (Ericson, Real-Time Collision Detection, 5.1.5 point-triangle and 5.1.9
  segment-segment), written independently below in numpy - the reference's
  own distance test uses 1e-6 (ptdistance/test.cpp:11-31).
* Gradients / Hessians agree with torch float64 autograd of the same
  squared-distance formula, and the mollifier with autograd of its closed form.
* Closest-point barycentrics reproduce the distance (the assertion of
  model/pt_ipc.cpp:67 and model/ee_ipc.cpp:94) and tangent bases are
  orthonormal and perpendicular to the contact normal.
"""
import numpy as np
import pytest
import torch

import oracle_lib as O
from adb_amd.scenes import PCG32

P_T0, P_T1, P_T2, P_E0, P_E1, P_E2, P_T = range(7)
EA0_EB0, EA0_EB1, EA1_EB0, EA1_EB1, EA_EB0, EA_EB1, EA0_EB, EA1_EB, EA_EB = range(9)


def closest_pt_triangle(p, a, b, c):
    """Ericson RTCD 5.1.5 ClosestPtPointTriangle."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + d1 / (d1 - d3) * ab
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return c
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + d2 / (d2 - d6) * ac
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (d4 - d3) / ((d4 - d3) + (d5 - d6)) * (c - b)
    denom = 1.0 / (va + vb + vc)
    return a + ab * (vb * denom) + ac * (vc * denom)


def closest_seg_seg(p1, q1, p2, q2):
    """Ericson RTCD 5.1.9 ClosestPtSegmentSegment -> squared distance."""
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, e, f = d1 @ d1, d2 @ d2, d2 @ r
    c = d1 @ r
    b = d1 @ d2
    denom = a * e - b * b
    s = np.clip((b * f - c * e) / denom, 0, 1) if denom != 0 else 0.0
    t = (b * s + f) / e
    if t < 0:
        t, s = 0.0, np.clip(-c / a, 0, 1)
    elif t > 1:
        t, s = 1.0, np.clip((b - c) / a, 0, 1)
    c1, c2 = p1 + d1 * s, p2 + d2 * t
    return float((c1 - c2) @ (c1 - c2))


def rand_pts(rng, n=4, scale=1.0):
    return np.array([[rng.uniform(-scale, scale) for _ in range(3)] for _ in range(n)])


def test_pt_distance_matches_textbook_closest_point():
    rng = PCG32(21)
    types = set()
    for _ in range(3000):
        P = rand_pts(rng)
        t, d = O.pt_distance(P.reshape(-1))
        c = closest_pt_triangle(P[0], P[1], P[2], P[3])
        ref = float((P[0] - c) @ (P[0] - c))
        assert d == pytest.approx(ref, rel=1e-9, abs=1e-15)
        types.add(t)
    assert types == set(range(7))


def test_ee_distance_matches_textbook_closest_points():
    rng = PCG32(22)
    types = set()
    for _ in range(3000):
        P = rand_pts(rng)
        t, d = O.ee_distance(P.reshape(-1))
        ref = closest_seg_seg(P[0], P[1], P[2], P[3])
        assert d == pytest.approx(ref, rel=1e-9, abs=1e-15)
        types.add(t)
    assert types == set(range(9))


def test_ee_parallel_edges():
    # exactly parallel, axis-aligned edges (the cube-stack case): distance is an
    # endpoint-to-segment distance, never the degenerate line-line formula
    cases = [
        ([0, 0, 0], [1, 0, 0], [0.2, 0.1, 0], [0.7, 0.1, 0], 0.01),       # overlapping spans
        ([0, 0, 0], [1, 0, 0], [1.5, 0.2, 0], [2.5, 0.2, 0], 0.25 + 0.04),  # disjoint spans
        ([0, 0, 0], [1, 0, 0], [-1, 0, 0.3], [0.5, 0, 0.3], 0.09),
    ]
    for a0, a1, b0, b1, ref in cases:
        P = np.array([a0, a1, b0, b1], dtype=float)
        t, d = O.ee_distance(P.reshape(-1))
        assert t != EA_EB
        assert d == pytest.approx(ref, rel=1e-12)
        assert d == pytest.approx(closest_seg_seg(*P), rel=1e-12)


# ---- derivatives vs torch autograd -------------------------------------------------
def _t_pp(a, b):
    e = a - b
    return (e * e).sum()


def _t_pl(p, e0, e1):
    c = torch.linalg.cross(e0 - p, e1 - p)
    e = e1 - e0
    return (c * c).sum() / (e * e).sum()


def _t_pplane(p, t0, t1, t2):
    n = torch.linalg.cross(t1 - t0, t2 - t0)
    return ((p - t0) * n).sum() ** 2 / (n * n).sum()


def _t_ll(a0, a1, b0, b1):
    n = torch.linalg.cross(a1 - a0, b1 - b0)
    return ((b0 - a0) * n).sum() ** 2 / (n * n).sum()


PT_FN = {P_T0: lambda P: _t_pp(P[0], P[1]), P_T1: lambda P: _t_pp(P[0], P[2]), P_T2: lambda P: _t_pp(P[0], P[3]),
         P_E0: lambda P: _t_pl(P[0], P[1], P[2]), P_E1: lambda P: _t_pl(P[0], P[2], P[3]),
         P_E2: lambda P: _t_pl(P[0], P[3], P[1]), P_T: lambda P: _t_pplane(P[0], P[1], P[2], P[3])}
EE_FN = {EA0_EB0: lambda P: _t_pp(P[0], P[2]), EA0_EB1: lambda P: _t_pp(P[0], P[3]),
         EA1_EB0: lambda P: _t_pp(P[1], P[2]), EA1_EB1: lambda P: _t_pp(P[1], P[3]),
         EA_EB0: lambda P: _t_pl(P[2], P[0], P[1]), EA_EB1: lambda P: _t_pl(P[3], P[0], P[1]),
         EA0_EB: lambda P: _t_pl(P[0], P[2], P[3]), EA1_EB: lambda P: _t_pl(P[1], P[2], P[3]),
         EA_EB: lambda P: _t_ll(P[0], P[1], P[2], P[3])}


def _autograd(fn, x):
    xt = torch.tensor(x, dtype=torch.float64)
    f = lambda v: fn(v.reshape(4, 3))
    g = torch.autograd.functional.jacobian(f, xt).numpy()
    H = torch.autograd.functional.hessian(f, xt).numpy()
    return g, H


@pytest.mark.parametrize("t", list(range(7)))
def test_pt_distance_derivatives_autograd(t):
    rng = PCG32(100 + t)
    for _ in range(6):
        x = rand_pts(rng).reshape(-1)
        g, H = O.pt_distance_derivs(x, t)
        gr, Hr = _autograd(PT_FN[t], x)
        np.testing.assert_allclose(g, gr, rtol=1e-10, atol=1e-12 * np.abs(gr).max())
        np.testing.assert_allclose(H, Hr, rtol=1e-9, atol=1e-11 * np.abs(Hr).max())


@pytest.mark.parametrize("t", list(range(9)))
def test_ee_distance_derivatives_autograd(t):
    rng = PCG32(200 + t)
    for _ in range(6):
        x = rand_pts(rng).reshape(-1)
        g, H = O.ee_distance_derivs(x, t)
        gr, Hr = _autograd(EE_FN[t], x)
        np.testing.assert_allclose(g, gr, rtol=1e-10, atol=1e-12 * np.abs(gr).max())
        np.testing.assert_allclose(H, Hr, rtol=1e-9, atol=1e-11 * np.abs(Hr).max())


def test_mollifier_autograd():
    rng = PCG32(300)
    for _ in range(10):
        P = rand_pts(rng)
        # make the edges nearly parallel so the mollifier is active
        P[3] = P[2] + (P[1] - P[0]) * 0.7 + 1e-2 * rand_pts(rng, 1)[0]
        x = P.reshape(-1)
        ea2 = float(((P[0] - P[1]) ** 2).sum())
        eb2 = float(((P[2] - P[3]) ** 2).sum())
        eps = 1e-3 * ea2 * eb2
        m, g, H = O.ee_mollifier(x, eps)

        def f(v):
            Q = v.reshape(4, 3)
            c = torch.linalg.cross(Q[1] - Q[0], Q[3] - Q[2])
            xx = (c * c).sum()
            return (-xx / eps + 2) * xx / eps

        xt = torch.tensor(x)
        xx = float(np.sum(np.cross(P[1] - P[0], P[3] - P[2]) ** 2))
        if xx < eps:
            assert m == pytest.approx(float(f(xt)), rel=1e-12)
            np.testing.assert_allclose(g, torch.autograd.functional.jacobian(f, xt).numpy(), rtol=1e-9, atol=1e-12)
            np.testing.assert_allclose(H, torch.autograd.functional.hessian(f, xt).numpy(), rtol=1e-8, atol=1e-10)
        else:
            assert m == 1.0 and np.all(g == 0) and np.all(H == 0)


def test_closest_points_reproduce_distance():
    rng = PCG32(31)
    for _ in range(500):
        P = rand_pts(rng)
        t, d = O.pt_distance(P.reshape(-1))
        lam, Pk = O.pt_closest(P.reshape(-1), t)
        c = lam[0] * P[1] + lam[1] * P[2] + lam[2] * P[3]
        assert float((P[0] - c) @ (P[0] - c)) == pytest.approx(d, rel=1e-9, abs=1e-14)
        np.testing.assert_allclose(Pk.T @ Pk, np.eye(2), atol=1e-12)
        n = P[0] - c
        if np.linalg.norm(n) > 1e-9:
            assert np.abs(Pk.T @ n).max() < 1e-9 * np.linalg.norm(n)
        t2, d2 = O.ee_distance(P.reshape(-1))
        lm, Pk2 = O.ee_closest(P.reshape(-1), t2)
        ca = (1 - lm[0]) * P[0] + lm[0] * P[1]
        cb = (1 - lm[1]) * P[2] + lm[1] * P[3]
        assert float((ca - cb) @ (ca - cb)) == pytest.approx(d2, rel=1e-8, abs=1e-14)
        np.testing.assert_allclose(Pk2.T @ Pk2, np.eye(2), atol=1e-12)
        n2 = ca - cb
        if np.linalg.norm(n2) > 1e-9:
            assert np.abs(Pk2.T @ n2).max() < 1e-8 * np.linalg.norm(n2)
