"""Pins for the oracle's contact compositions that the unit pins leave open.

* Mollified edge-edge barrier (model/core_ipc.cpp:47-83): the oracle's
  product-rule Hessian and gradient of m(x; eps) B(d(x)) against torch float64
  autograd of that product, on near-parallel edge pairs where m < 1 (eps held
  at its value from the base configuration, as the reference treats it).
* Point-triangle and edge-edge lagged friction (model/pt_ipc.cpp:22-107,
  model/ee_ipc.cpp:21-118, model/friction.cpp:15-47, model/energy.cpp:195-258):
  the friction part of the assembled gradient and Hessian at a sliding Newton
  iterate, and the friction energy, against torch autograd of
  sum_k mu lambda_k f0(|(I - n n^T)(sum_i w_i (x_i - x_i^t0))|), where the
  closest-point weights w and the normal n come from the textbook closest-point
  routines (Ericson, Real-Time Collision Detection 5.1.5 and 5.1.9) and
  lambda_k = -B'(d) 2 sqrt(d) from the closed-form barrier: every tangent basis
  of the reference spans the plane orthogonal to the closest-point direction,
  so |Tk (x - x_t0)| is this projected norm whatever basis is chosen.
* A cube sliding on a static cube decelerates at mu g (Coulomb), once with
  point-triangle contacts (a smaller cube on a larger one) and once with
  crossing edge-edge contacts (equal cubes rotated by 45 degrees).
* The line search (model/line_search.cpp:39-136) returns the first halving
  that satisfies the Armijo condition.
"""
import math

import numpy as np
import pytest
import torch

import oracle_lib as O
from adb_amd import scenes
from adb_amd.capi import default_params, kernel_config
from adb_amd.scenes import PCG32


# ---- closed forms (model/barrier.cpp:11-28, model/FrictionUtils.h:5-32) ------
def t_barrier(d, dh, kap):
    return kap * -(d - dh) * (d - dh) * torch.log(d / dh) / (dh * dh)


def barrier_d1(d, dh, kap):
    return -(d - dh) * kap * (2 * math.log(d / dh) + (d - dh) / d) / (dh * dh)


def t_f0(u, eps):
    if float(u.detach()) >= eps:
        return u
    return u * u * (-u / 3 + eps) / (eps * eps) + eps / 3


# ---- textbook distances in torch (squared) -----------------------------------
def t_pp(a, b):
    e = a - b
    return (e * e).sum()


def t_pl(p, e0, e1):
    c = torch.linalg.cross(e0 - p, e1 - p)
    e = e1 - e0
    return (c * c).sum() / (e * e).sum()


def t_ll(a0, a1, b0, b1):
    n = torch.linalg.cross(a1 - a0, b1 - b0)
    return ((b0 - a0) * n).sum() ** 2 / (n * n).sum()


EE_FN = [lambda P: t_pp(P[0], P[2]), lambda P: t_pp(P[0], P[3]), lambda P: t_pp(P[1], P[2]),
         lambda P: t_pp(P[1], P[3]), lambda P: t_pl(P[2], P[0], P[1]), lambda P: t_pl(P[3], P[0], P[1]),
         lambda P: t_pl(P[0], P[2], P[3]), lambda P: t_pl(P[1], P[2], P[3]), lambda P: t_ll(P[0], P[1], P[2], P[3])]


def t_mollifier(P, eps):
    c = torch.linalg.cross(P[1] - P[0], P[3] - P[2])
    x = (c * c).sum()
    if float(x.detach()) >= eps:
        return torch.ones((), dtype=torch.float64)
    return (-x / eps + 2) * x / eps


def near_parallel_pairs(rng, n):
    """Edge pairs with |ea x eb|^2 < 1e-3 |ea|^2 |eb|^2 (mollifier active) and a
    squared distance below 1e-4 (the abdtk d_hat)."""
    out = []
    while len(out) < n:
        a0 = np.array([rng.uniform(-1, 1) for _ in range(3)])
        u = np.array([rng.normal() for _ in range(3)]); u /= np.linalg.norm(u)
        w = np.cross(u, [rng.normal() for _ in range(3)]); w /= np.linalg.norm(w)
        v = np.cross(u, w)
        la, lb = rng.uniform(0.5, 1.5), rng.uniform(0.5, 1.5)
        th = rng.uniform(0.002, 0.025)
        eb = math.cos(th) * u + math.sin(th) * v
        gap = rng.uniform(0.002, 0.008)
        shift = rng.uniform(-0.6, 0.6) * la
        b0 = a0 + shift * u + gap * w
        P = np.array([a0, a0 + la * u, b0, b0 + lb * eb])
        t, d = O.ee_distance(P.reshape(-1))
        eps = 1e-3 * la * la * lb * lb
        cr = np.cross(P[1] - P[0], P[3] - P[2])
        if d < 1e-4 and cr @ cr < eps:
            out.append((P, t, d, eps))
    return out


def test_mollified_ee_matches_autograd():
    rng = PCG32(505)
    cfg = kernel_config(d_hat=1e-4, kappa=1e-3, eps_x=1e-3, psd=0)
    types = set()
    for P, t, d, eps in near_parallel_pairs(rng, 40):
        H, g, m = O.ipc_hess_ee(cfg, P.reshape(-1), t, d)
        assert 0.0 < m < 1.0

        def energy(v):
            Q = v.reshape(4, 3)
            return t_mollifier(Q, eps) * t_barrier(EE_FN[t](Q), 1e-4, 1e-3)

        x = torch.tensor(P.reshape(-1), dtype=torch.float64)
        gr = torch.autograd.functional.jacobian(energy, x).numpy()
        Hr = torch.autograd.functional.hessian(energy, x).numpy()
        np.testing.assert_allclose(g, gr, rtol=1e-8, atol=1e-10 * np.abs(gr).max())
        np.testing.assert_allclose(H, Hr, rtol=1e-7, atol=1e-9 * np.abs(Hr).max())
        types.add(t)
    # both the interior (line-line) and the endpoint types of the parallel rule
    assert 8 in types and len(types) >= 2


def test_mollifier_product_rule_is_not_plain_product():
    # planted mistakes fail: dropping the cross terms B'(grad m grad d^T + ...)
    # or the m B' hess d term moves the Hessian far beyond the tolerance above
    rng = PCG32(506)
    cfg = kernel_config(d_hat=1e-4, kappa=1e-3, eps_x=1e-3, psd=0)
    P, t, d, eps = near_parallel_pairs(rng, 1)[0]
    H, g, m = O.ipc_hess_ee(cfg, P.reshape(-1), t, d)
    gd, Hd = O.ee_distance_derivs(P.reshape(-1), t)
    mm, gm, Hm = O.ee_mollifier(P.reshape(-1), eps)
    B0, B1, B2 = O.barrier(d, 1e-4, 1e-3)
    full = Hm * B0 + B1 * (np.outer(gm, gd) + np.outer(gd, gm)) + m * (B2 * np.outer(gd, gd) + B1 * Hd)
    np.testing.assert_allclose(H, full, rtol=1e-12, atol=1e-12 * np.abs(full).max())
    for wrong in (Hm * B0 + m * (B2 * np.outer(gd, gd) + B1 * Hd),
                  Hm * B0 + B1 * (np.outer(gm, gd) + np.outer(gd, gm)) + m * B2 * np.outer(gd, gd)):
        assert np.abs(wrong - H).max() > 1e-6 * np.abs(H).max()


# ---- textbook closest points with weights ------------------------------------
def closest_pt_triangle_weights(p, a, b, c):
    """Ericson RTCD 5.1.5, returning the barycentric weights of a, b, c."""
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
    v, w = vb * den, vc * den
    return np.array([1 - v - w, v, w])


def closest_seg_seg_params(p1, q1, p2, q2):
    """Ericson RTCD 5.1.9: parameters (s, t) of the closest points."""
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


def world_x(sc, q, v):
    b = int(np.searchsorted(sc["vert_offset"], v, side="right") - 1)
    A = q[b, 3:].reshape(3, 3).T
    return A @ sc["rest_verts"][v] + q[b, :3]


def t_world_x(sc, qt, v):
    b = int(np.searchsorted(sc["vert_offset"], v, side="right") - 1)
    xb = torch.tensor(sc["rest_verts"][v], dtype=torch.float64)
    qb = qt[12 * b:12 * b + 12]
    return qb[0:3] + qb[3:6] * xb[0] + qb[6:9] * xb[1] + qb[9:12] * xb[2]


def friction_terms(sc, q, q0, pt, ee, d_hat, kappa):
    """Per active contact: global vertex ids, textbook weights, unit normal,
    lambda (all lagged at q)."""
    vo, fo, eo = sc["vert_offset"], sc["face_offset"], sc["edge_offset"]
    out = []
    for I, v, J, f in pt:
        ids = [vo[I] + v] + [vo[J] + k for k in sc["faces"][fo[J] + f]]
        P = np.array([world_x(sc, q, i) for i in ids])
        w3 = closest_pt_triangle_weights(P[0], P[1], P[2], P[3])
        c = w3 @ P[1:]
        d = float((P[0] - c) @ (P[0] - c))
        if d >= d_hat:
            continue
        out.append((ids, np.concatenate([[-1.0], w3]), (P[0] - c) / math.sqrt(d), -barrier_d1(d, d_hat, kappa) * 2 * math.sqrt(d)))
    for I, a, J, b in ee:
        ids = [vo[I] + k for k in sc["edges"][eo[I] + a]] + [vo[J] + k for k in sc["edges"][eo[J] + b]]
        P = np.array([world_x(sc, q, i) for i in ids])
        s, t = closest_seg_seg_params(P[0], P[1], P[2], P[3])
        ca, cb = (1 - s) * P[0] + s * P[1], (1 - t) * P[2] + t * P[3]
        d = float((ca - cb) @ (ca - cb))
        if d >= d_hat:
            continue
        out.append((ids, np.array([-(1 - s), -s, 1 - t, t]), (ca - cb) / math.sqrt(d), -barrier_d1(d, d_hat, kappa) * 2 * math.sqrt(d)))
    return out


def textbook_friction_energy(sc, terms, q0, mu, eps):
    def energy(qt):
        e = torch.zeros((), dtype=torch.float64)
        for ids, w, n, lam in terms:
            rel = torch.zeros(3, dtype=torch.float64)
            for i, wi in zip(ids, w):
                rel = rel + wi * (t_world_x(sc, qt, i) - torch.tensor(world_x(sc, q0, i)))
            nt = torch.tensor(n)
            u = rel - nt * (nt @ rel)
            e = e + mu * lam * t_f0(torch.sqrt(u @ u), eps)
        return e
    return energy


def stacked_pair(top_size, angle_deg, gap):
    """A static unit cube and a dynamic cube of edge top_size rotated about y
    resting `gap` above it."""
    v0, f0, e0 = scenes.cube_mesh()
    v1, f1, e1 = scenes.cube_mesh((top_size, top_size, top_size))
    R = np.array([[math.cos(math.radians(angle_deg)), 0, math.sin(math.radians(angle_deg))], [0, 1, 0],
                  [-math.sin(math.radians(angle_deg)), 0, math.cos(math.radians(angle_deg))]])
    return scenes.build_scene([dict(verts=v0, tris=f0, edges=e0, p=(0.0, 0.0, 0.0), mass=-1000.0, Ic=1000.0 / 12),
                               dict(verts=v1, tris=f1, edges=e1, R=R, p=(0.05, 0.5 + top_size / 2 + gap, -0.03))])


@pytest.mark.parametrize("top_size,angle", [(0.5, 30.0), (1.0, 45.0)])
def test_pt_ee_friction_matches_textbook(top_size, angle):
    d_hat, kappa, mu, dt, evh = 1e-3, 1e-4, 0.5, 0.01, 1e-3
    sc = stacked_pair(top_size, angle, 0.5 * math.sqrt(d_hat))
    q0 = sc["q0"].copy()
    q = q0.copy()
    q[1, 0] += 3e-3          # slide the top cube (u >> eps = dt evh)
    q[1, 2] -= 1.5e-3
    worlds = {m: O.OracleWorld(sc, default_params(d_hat=d_hat, kappa=kappa, mu=m, dt=dt, evh=evh, enable_ground=0))
              for m in (0.0, mu)}
    g1, H1 = worlds[mu].assemble_at(q)
    g0, H0 = worlds[0.0].assemble_at(q)
    pt, ee, _ = worlds[mu].collision_set(vtn=2, dq=q - q0)
    terms = friction_terms(sc, q, q0, pt, ee, d_hat, kappa)
    assert len(terms) >= 4
    energy = textbook_friction_energy(sc, terms, q0, mu, dt * evh)
    qt = torch.tensor(q.reshape(-1), dtype=torch.float64)
    gr = torch.autograd.functional.jacobian(energy, qt).numpy()
    Hr = torch.autograd.functional.hessian(energy, qt).numpy()
    # the static body's rows are not assembled (model/assemble_matrices.cpp:43-58)
    sl = slice(12, 24)
    np.testing.assert_allclose((g1 - g0)[sl], gr[sl], rtol=1e-7, atol=1e-9 * np.abs(gr[sl]).max())
    np.testing.assert_allclose((H1 - H0)[sl, sl], Hr[sl, sl], rtol=1e-6, atol=1e-8 * np.abs(Hr[sl, sl]).max())
    assert worlds[mu].friction_energy(q) == pytest.approx(float(energy(qt)), rel=1e-10)
    # point-triangle contacts in the small-cube case, crossing edge pairs in both
    assert (len(pt) > 0) == (top_size < 1.0) and len(ee) > 0


@pytest.mark.parametrize("top_size,angle", [(0.5, 30.0), (1.0, 45.0)])
def test_friction_textbook_check_catches_planted_weight_error(top_size, angle):
    # the check above is sensitive: a wrong barycentric weight changes the
    # gradient far beyond its tolerance
    d_hat, kappa, mu, dt, evh = 1e-3, 1e-4, 0.5, 0.01, 1e-3
    sc = stacked_pair(top_size, angle, 0.5 * math.sqrt(d_hat))
    q0 = sc["q0"].copy()
    q = q0.copy()
    q[1, 0] += 3e-3
    q[1, 2] -= 1.5e-3
    ow = O.OracleWorld(sc, default_params(d_hat=d_hat, kappa=kappa, mu=mu, dt=dt, evh=evh, enable_ground=0))
    pt, ee, _ = ow.collision_set(vtn=2, dq=q - q0)
    terms = friction_terms(sc, q, q0, pt, ee, d_hat, kappa)
    bad = [(ids, w[::-1].copy(), n, lam) for ids, w, n, lam in terms]
    qt = torch.tensor(q.reshape(-1), dtype=torch.float64)
    g = torch.autograd.functional.jacobian(textbook_friction_energy(sc, terms, q0, mu, dt * evh), qt).numpy()
    gb = torch.autograd.functional.jacobian(textbook_friction_energy(sc, bad, q0, mu, dt * evh), qt).numpy()
    assert np.abs(g - gb)[12:].max() > 1e-3 * np.abs(g[12:]).max()


@pytest.mark.parametrize("top_size,angle", [(0.5, 30.0), (1.0, 45.0)])
def test_cube_sliding_on_cube_decelerates_at_mu_g(top_size, angle):
    mu, g, dt = 0.5, 9.8, 0.01
    sc = stacked_pair(top_size, angle, 0.5 * math.sqrt(1e-3))
    p = default_params(d_hat=1e-3, mu=mu, tol=1e-6, enable_ground=0)
    w = O.OracleWorld(sc, p)
    for _ in range(40):        # settle onto the barrier
        assert w.step()[0] == 0
    q, qd = w.state()
    qd[1, :3] = (0.8, 0.0, 0.6)   # |v| = 1
    w.set_state(q, qd)
    n = 10
    for _ in range(n):
        rc, st = w.step()
        assert rc == 0
        assert st.n_pt + st.n_ee > 0 and st.n_vg == 0
    q, qd = w.state()
    speed = float(np.linalg.norm(qd[1, :3]))
    assert speed == pytest.approx(1.0 - mu * g * dt * n, rel=0.03)
    # the velocity keeps its direction up to the yaw that the off-centre contact
    # patch's friction torque gives the body (isotropic friction per contact)
    assert np.dot(qd[1, [0, 2]], [0.8, 0.6]) / speed == pytest.approx(1.0, abs=1e-2)


def test_line_search_returns_first_armijo_halving():
    sc = scenes.cube_stack(3, seed=2)
    p = default_params(d_hat=1e-3, mu=0.5)
    w = O.OracleWorld(sc, p)
    q = sc["q0"].copy()
    g, _ = w.assemble_at(q)
    rng = np.random.default_rng(5)
    for scale in (1e-4, 1e-2, 0.3):
        dq = -g * scale / np.abs(g).max() + rng.normal(0, 1e-3 * scale, size=g.shape)
        rc, alpha, E0, E1, trials = w.line_search(q, dq)
        assert rc == 0
        assert alpha == 0.5 ** (trials - 1)
        qdg = float(dq @ g)
        assert E1 <= E0 + 1e-4 * alpha * qdg
        if trials > 1:
            # the previous (twice as long) trial violated the condition
            w.assemble_at(q)
            Eprev = w.energy(q.reshape(-1) + 2 * alpha * dq) + w.friction_energy(q.reshape(-1) + 2 * alpha * dq)
            assert Eprev > E0 + 1e-4 * 2 * alpha * qdg
