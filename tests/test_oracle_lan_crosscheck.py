"""Cross-check of the body-space contact terms, as the reference's AIPC ("lan")
comparison does it (model/lan_friction.cpp:14-203: per contact the gradients
ga, gb and Hessian blocks ha, hb, hab of the two bodies from the AIPC
constraint classes, model/IpcCollisionConstraint.cpp and
model/IpcFrictionConstraint.cpp, compared with ipc_term's gradp, gradt,
hess_p, hess_t, off_diag).

Here the second formulation is written independently in torch: the contact
energy as a function of the 24 affine coordinates of the two bodies,
x_k = p + [a1 a2 a3] xbar_k (the AIPC weights w = (1, xbar)), the barrier of
model/barrier.cpp:11-28 on the squared distance of the contact's type (times
the edge-edge mollifier with eps_x |ea|^2 |eb|^2 held at its current value),
plus the lagged friction potential mu lambda f0(|(I - n n^T) sum_i w_i (x_i -
x_i^t0)|), differentiated by autograd.  The oracle's assembled blocks (psd off,
inertia removed by assembling the same scene without contacts) must agree
within 1e-8 of the largest entry, for every point-triangle and edge-edge
distance type, mollified edge pairs included.
"""
import math

import numpy as np
import pytest
import torch

import oracle_lib as O
from adb_amd import scenes
from adb_amd.capi import default_params

D_HAT, KAPPA, MU, DT, EVH = 1e-3, 1e-4, 0.4, 0.01, 1e-3


def t_barrier(d):
    return KAPPA * -(d - D_HAT) * (d - D_HAT) * torch.log(d / D_HAT) / (D_HAT * D_HAT)


def barrier_d1(d):
    return -(d - D_HAT) * KAPPA * (2 * math.log(d / D_HAT) + (d - D_HAT) / d) / (D_HAT * D_HAT)


def t_pp(a, b):
    e = a - b
    return (e * e).sum()


def t_pl(p, e0, e1):
    c = torch.linalg.cross(e0 - p, e1 - p)
    e = e1 - e0
    return (c * c).sum() / (e * e).sum()


def t_pplane(p, t0, t1, t2):
    n = torch.linalg.cross(t1 - t0, t2 - t0)
    return ((p - t0) * n).sum() ** 2 / (n * n).sum()


def t_ll(a0, a1, b0, b1):
    n = torch.linalg.cross(a1 - a0, b1 - b0)
    return ((b0 - a0) * n).sum() ** 2 / (n * n).sum()


PT_FN = [lambda P: t_pp(P[0], P[1]), lambda P: t_pp(P[0], P[2]), lambda P: t_pp(P[0], P[3]),
         lambda P: t_pl(P[0], P[1], P[2]), lambda P: t_pl(P[0], P[2], P[3]), lambda P: t_pl(P[0], P[3], P[1]),
         lambda P: t_pplane(P[0], P[1], P[2], P[3])]
EE_FN = [lambda P: t_pp(P[0], P[2]), lambda P: t_pp(P[0], P[3]), lambda P: t_pp(P[1], P[2]),
         lambda P: t_pp(P[1], P[3]), lambda P: t_pl(P[2], P[0], P[1]), lambda P: t_pl(P[3], P[0], P[1]),
         lambda P: t_pl(P[0], P[2], P[3]), lambda P: t_pl(P[1], P[2], P[3]), lambda P: t_ll(P[0], P[1], P[2], P[3])]


def body(verts, tris=(), edges=()):
    return dict(verts=np.asarray(verts, dtype=float), tris=np.asarray(tris, dtype=np.int32).reshape(-1, 3),
                edges=np.asarray(edges, dtype=np.int32).reshape(-1, 2), p=(0.0, 0.0, 0.0), mass=1.0, Ic=1.0)


def affine(rng, scale=0.05):
    A = np.eye(3) + rng.normal(0, scale, size=(3, 3))
    return np.concatenate([rng.normal(0, 0.002, size=3), A[:, 0], A[:, 1], A[:, 2]])


def world(sc, q):
    out = []
    for b in range(len(sc["mass"])):
        v0, v1 = sc["vert_offset"][b], sc["vert_offset"][b + 1]
        A = q[b, 3:].reshape(3, 3).T
        out.append(sc["rest_verts"][v0:v1] @ A.T + q[b, :3])
    return np.concatenate(out)


def closest_pt_triangle_weights(p, a, b, c):
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return np.array([1.0, 0.0, 0.0])
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return np.array([0.0, 1.0, 0.0])
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        v = d1 / (d1 - d3)
        return np.array([1 - v, v, 0.0])
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return np.array([0.0, 0.0, 1.0])
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        w = d2 / (d2 - d6)
        return np.array([1 - w, 0.0, w])
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        w = (d4 - d3) / ((d4 - d3) + (d5 - d6))
        return np.array([0.0, 1 - w, w])
    den = 1.0 / (va + vb + vc)
    return np.array([1 - vb * den - vc * den, vb * den, vc * den])


def closest_seg_seg_params(p1, q1, p2, q2):
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, e, f = d1 @ d1, d2 @ d2, d2 @ r
    c, b = d1 @ r, d1 @ d2
    den = a * e - b * b
    s = float(np.clip((b * f - c * e) / den, 0, 1)) if den != 0 else 0.0
    t = (b * s + f) / e
    if t < 0:
        t, s = 0.0, float(np.clip(-c / a, 0, 1))
    elif t > 1:
        t, s = 1.0, float(np.clip((b - c) / a, 0, 1))
    return s, t


def blocks_of(sc, q, q0, kind, with_friction):
    """Oracle blocks (contact part): assembled with contacts minus without."""
    mu = MU if with_friction else 0.0
    on = default_params(d_hat=D_HAT, kappa=KAPPA, mu=mu, dt=DT, evh=EVH, psd=0, enable_ground=0)
    off = default_params(d_hat=D_HAT, kappa=KAPPA, mu=mu, dt=DT, evh=EVH, psd=0, enable_ground=0,
                         enable_pt=0, enable_ee=0)
    g1, H1 = O.OracleWorld(sc, on).assemble_at(q)
    g0, H0 = O.OracleWorld(sc, off).assemble_at(q)
    return g1 - g0, H1 - H0


def lan_terms(sc, q, q0, kind, ty, eps, with_friction):
    """Gradient and Hessian over the 24 body coordinates by autograd."""
    xb = torch.tensor(sc["rest_verts"])
    vo = sc["vert_offset"]

    def pts(qq):
        Q = qq.reshape(2, 12)
        out = []
        for b in range(2):
            A = torch.stack([Q[b, 3:6], Q[b, 6:9], Q[b, 9:12]], dim=1)
            for v in range(vo[b], vo[b + 1]):
                out.append(A @ xb[v] + Q[b, :3])
        return out

    X = world(sc, q)
    Xs = world(sc, q0)
    if kind == "pt":
        ids = [0, 1, 2, 3]                    # the point, then the triangle
        w3 = closest_pt_triangle_weights(X[0], X[1], X[2], X[3])
        wts = np.concatenate([[-1.0], w3])
        c = w3 @ X[1:4]
        n = (X[0] - c) / np.linalg.norm(X[0] - c)
    else:
        ids = [0, 1, 2, 3]
        s, t = closest_seg_seg_params(X[0], X[1], X[2], X[3])
        wts = np.array([-(1 - s), -s, 1 - t, t])
        ca, cb = (1 - s) * X[0] + s * X[1], (1 - t) * X[2] + t * X[3]
        n = (ca - cb) / np.linalg.norm(ca - cb)
    d0 = float((PT_FN if kind == "pt" else EE_FN)[ty](torch.tensor(X[ids])))
    lam = -barrier_d1(d0) * 2 * math.sqrt(d0)

    def energy(qq):
        P = pts(qq)
        P4 = torch.stack([P[i] for i in ids])
        d = (PT_FN if kind == "pt" else EE_FN)[ty](P4)
        e = t_barrier(d)
        if kind == "ee":
            cr = torch.linalg.cross(P4[1] - P4[0], P4[3] - P4[2])
            xx = (cr * cr).sum()
            if float(xx.detach()) < eps:
                e = e * (-xx / eps + 2) * xx / eps
        if with_friction:
            rel = torch.zeros(3, dtype=torch.float64)
            for i, wi in zip(ids, wts):
                rel = rel + wi * (P[i] - torch.tensor(Xs[i]))
            nt = torch.tensor(n)
            u = rel - nt * (nt @ rel)
            un = torch.sqrt(u @ u)
            ev = DT * EVH
            f0 = un if float(un.detach()) >= ev else un * un * (-un / 3 + ev) / (ev * ev) + ev / 3
            e = e + MU * lam * f0
        return e

    qt = torch.tensor(q.reshape(-1))
    return torch.autograd.functional.jacobian(energy, qt).numpy(), torch.autograd.functional.hessian(energy, qt).numpy()


def pt_case(rng, want_type):
    tri = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.1], [0.2, 0.0, 1.0]])
    tri += rng.normal(0, 0.05, size=(3, 3))
    # the point placed near the region of the wanted closest feature
    if want_type == 6:
        base = (tri[0] + tri[1] + tri[2]) / 3
    elif want_type < 3:
        v = tri[want_type]
        e1, e2 = tri[(want_type + 1) % 3] - v, tri[(want_type + 2) % 3] - v
        out = -(e1 / np.linalg.norm(e1) + e2 / np.linalg.norm(e2))      # outward bisector: the vertex region
        base = v + out / np.linalg.norm(out) * 0.01
    else:
        a, b = [(0, 1), (1, 2), (2, 0)][want_type - 3]
        out = 0.5 * (tri[a] + tri[b]) - tri.mean(0)
        base = 0.5 * (tri[a] + tri[b]) + out / np.linalg.norm(out) * 0.01
    n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
    n /= np.linalg.norm(n)
    pnt = base + n * rng.uniform(0.008, 0.02)
    sc = scenes.build_scene([body([pnt]), body(tri, [(0, 1, 2)], [(0, 1), (1, 2), (0, 2)])])
    return sc


def ee_case(rng, parallel):
    a0, a1 = np.array([-0.5, 0.0, 0.0]), np.array([0.5, 0.0, 0.0])
    ang = rng.uniform(0.005, 0.02) if parallel else rng.uniform(0.3, 1.4)
    off = rng.uniform(-0.6, 0.6) if parallel else rng.uniform(-0.3, 0.3)
    gap = rng.uniform(0.008, 0.02)
    dirb = np.array([math.cos(ang), 0.0, math.sin(ang)])
    b0 = np.array([off, gap, 0.0]) - 0.5 * dirb
    b1 = b0 + dirb
    return scenes.build_scene([body([a0, a1], edges=[(0, 1)]), body([b0, b1], edges=[(0, 1)])])


@pytest.mark.parametrize("with_friction", [False, True])
def test_point_triangle_all_types(with_friction):
    rng = np.random.default_rng(41)
    seen = set()
    for want in range(7):
        for rep in range(4):
            sc = pt_case(rng, want)
            ident = np.concatenate([[0, 0, 0], np.eye(3).T.reshape(-1)])
            q0 = np.array([ident, ident])
            q = np.array([affine(rng, 0.002) for _ in range(2)])
            sc["q0"] = q0.copy()
            X = world(sc, q)
            ty, d = O.pt_distance(X[:4].reshape(-1))
            if not d < D_HAT:
                continue
            g, H = blocks_of(sc, q, q0, "pt", with_friction)
            gr, Hr = lan_terms(sc, q, q0, "pt", ty, 0.0, with_friction)
            np.testing.assert_allclose(g, gr, rtol=0, atol=1e-8 * np.abs(gr).max())
            np.testing.assert_allclose(H, Hr, rtol=0, atol=1e-8 * np.abs(Hr).max())
            seen.add(ty)
    assert seen == set(range(7)), seen


@pytest.mark.parametrize("with_friction", [False, True])
def test_edge_edge_types_and_mollifier(with_friction):
    rng = np.random.default_rng(42)
    seen, mollified = set(), 0
    for rep in range(24):
        sc = ee_case(rng, parallel=rep % 3 == 0)
        q0 = np.concatenate([[0, 0, 0], np.eye(3).T.reshape(-1)])
        sc["q0"] = np.array([q0, q0])
        q = np.array([affine(rng, 0.01) for _ in range(2)])
        X = world(sc, q)
        ty, d = O.ee_distance(X[:4].reshape(-1))
        if not d < D_HAT:
            continue
        ea, eb = X[1] - X[0], X[3] - X[2]
        eps = 1e-3 * (ea @ ea) * (eb @ eb)
        cr = np.cross(ea, eb)
        mollified += cr @ cr < eps
        g, H = blocks_of(sc, q, sc["q0"], "ee", with_friction)
        gr, Hr = lan_terms(sc, q, sc["q0"], "ee", ty, eps, with_friction)
        np.testing.assert_allclose(g, gr, rtol=0, atol=1e-8 * np.abs(gr).max())
        np.testing.assert_allclose(H, Hr, rtol=0, atol=1e-8 * np.abs(Hr).max())
        seen.add(ty)
    assert mollified >= 3 and 8 in seen and len(seen) >= 3, (mollified, seen)
