"""World-level pins for the oracle: contact sets, assembly, solve and short
trajectories with closed-form answers.

* Contact sets: the iAABB culling equals brute force over all primitive pairs
  (the `pipelined` test, iAABB/test.cpp:317-411, on its 50-random-cube fixture
  iAABB/test.cpp:89-169), and both agree with an independent numpy predicate.
* Assembly: with PSD projection, friction and ground off, the assembled
  gradient and Hessian are the exact derivatives of the incremental potential
  (inertia + orthogonality + barriers, model/energy.cpp:19-86); checked by
  finite differences through J = [I, x0 I, x1 I, x2 I] (model/barrier.cpp:42-54).
* Solve: dq = -H^{-1} g against a dense numpy solve.
* Trajectories: free fall is the implicit-Euler closed form; a cube settles on
  the ground at the gap where the VG barrier balances gravity; friction
  decelerates a sliding cube at about mu g.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from adb_amd import scenes
from adb_amd.capi import default_params
from test_oracle_distance import closest_pt_triangle, closest_seg_seg


def world_positions(sc, q):
    out = []
    for b in range(len(sc["mass"])):
        v0, v1 = sc["vert_offset"][b], sc["vert_offset"][b + 1]
        A = q[b, 3:].reshape(3, 3).T
        out.append(sc["rest_verts"][v0:v1] @ A.T + q[b, :3])
    return out


def numpy_sets(sc, q, d_hat, margin=1e-9):
    """Independent predicate with textbook distances; returns (surely-in, surely-out)
    decision functions for PT/EE keys."""
    X = world_positions(sc, q)
    B = len(sc["mass"])
    fo, eo = sc["face_offset"], sc["edge_offset"]
    pt_in, pt_border, ee_in, ee_border = set(), set(), set(), set()
    for I in range(B):
        for J in range(B):
            if I == J:
                continue
            for v in range(len(X[I])):
                for f in range(fo[J + 1] - fo[J]):
                    tri = sc["faces"][fo[J] + f]
                    c = closest_pt_triangle(X[I][v], X[J][tri[0]], X[J][tri[1]], X[J][tri[2]])
                    d = float((X[I][v] - c) @ (X[I][v] - c))
                    if d < d_hat * (1 - margin):
                        pt_in.add((I, v, J, f))
                    elif d < d_hat * (1 + margin):
                        pt_border.add((I, v, J, f))
            if I < J:
                for a in range(eo[I + 1] - eo[I]):
                    ea = sc["edges"][eo[I] + a]
                    for b in range(eo[J + 1] - eo[J]):
                        eb = sc["edges"][eo[J] + b]
                        d = closest_seg_seg(X[I][ea[0]], X[I][ea[1]], X[J][eb[0]], X[J][eb[1]])
                        if d < d_hat * (1 - margin):
                            ee_in.add((I, a, J, b))
                        elif d < d_hat * (1 + margin):
                            ee_border.add((I, a, J, b))
    return pt_in, pt_border, ee_in, ee_border


def keyset(a):
    return set(map(tuple, a.tolist()))


@pytest.mark.parametrize("seed", [7, 8])
def test_iaabb_equals_brute_force_random_cubes(seed):
    sc = scenes.random_cubes(50, seed=seed, box=2.0)
    p = default_params(d_hat=0.05)   # large threshold so the random fixture has contacts
    w = O.OracleWorld(sc, p)
    pt1, ee1, vg1 = w.collision_set(method=1)
    pt0, ee0, vg0 = w.collision_set(method=0)
    assert np.array_equal(pt1, pt0) and np.array_equal(ee1, ee0) and np.array_equal(vg1, vg0)
    assert len(pt1) > 0 and len(ee1) > 0
    # canonical order
    assert pt1.tolist() == sorted(pt1.tolist()) and ee1.tolist() == sorted(ee1.tolist())
    assert np.all(ee1[:, 0] < ee1[:, 2]) and np.all(pt1[:, 0] != pt1[:, 2])


def test_sets_match_independent_predicate():
    sc = scenes.random_cubes(12, seed=9, box=1.2)
    p = default_params(d_hat=0.04)
    w = O.OracleWorld(sc, p)
    pt, ee, vg = w.collision_set(method=1)
    pt_in, pt_b, ee_in, ee_b = numpy_sets(sc, sc["q0"], 0.04)
    assert len(pt_in) > 0 and len(ee_in) > 0
    got_pt, got_ee = keyset(pt), keyset(ee)
    assert pt_in <= got_pt and got_pt <= (pt_in | pt_b)
    assert ee_in <= got_ee and got_ee <= (ee_in | ee_b)
    # VG: (y - ground)^2 < d_hat
    X = world_positions(sc, sc["q0"])
    ref_vg = {(b, v) for b in range(len(X)) for v in range(len(X[b])) if (X[b][v][1] + 0.5) ** 2 < 0.04}
    assert keyset(vg) == ref_vg


def contact_scene():
    sc = scenes.cube_stack(3, seed=3, jitter=0.05, spacing=1.01, y0=0.5, rotate=False)
    # tilt bodies slightly so PT, EE and parallel-EE contacts all appear
    rng = scenes.PCG32(17)
    for b in range(3):
        R = scenes.euler_rotation(0.02 * rng.uniform(-1, 1), 0.02 * rng.uniform(-1, 1), 0.02 * rng.uniform(-1, 1))
        sc["q0"][b, 3:] = np.concatenate([R[:, 0], R[:, 1], R[:, 2]])
    sc["mass"][:] = 1.0
    sc["Ic"][:] = 1.0 / 12
    return sc


@pytest.mark.parametrize("seed", [7, 9])
def test_assembly_is_exact_derivative(seed):
    # Random orientations keep every EE pair away from the parallel regime, where
    # the reference's mollifier uses eps_x from the CURRENT edge lengths as a
    # constant (model/core_ipc.cpp:61) and so is not an exact derivative.
    sc = scenes.random_cubes(20, seed=seed, box=1.5)
    sc["mass"][:] = 1.0
    sc["Ic"][:] = 1.0 / 12
    p = default_params(d_hat=0.05, kappa=1e2, kappa_ortho=1.0, psd=0, mu=0.0, enable_ground=0)
    w = O.OracleWorld(sc, p)
    pt, ee, _ = w.collision_set()
    assert len(pt) > 0 and len(ee) > 0
    g, H = w.assemble(dense=True)
    q = sc["q0"].reshape(-1).copy()
    n = len(q)
    h = 1e-7
    gfd = np.zeros(n)
    for i in range(n):
        e = np.zeros(n); e[i] = h
        gfd[i] = (w.energy(q + e) - w.energy(q - e)) / (2 * h)
    np.testing.assert_allclose(g, gfd, rtol=1e-5, atol=1e-6 * np.abs(g).max())
    # Hessian: FD of the assembled gradient
    Hfd = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n); e[i] = h
        w.set_state((q + e).reshape(-1, 12))
        gp, _ = w.assemble(dense=False)
        w.set_state((q - e).reshape(-1, 12))
        gm, _ = w.assemble(dense=False)
        Hfd[:, i] = (gp - gm) / (2 * h)
    w.set_state(q.reshape(-1, 12))
    np.testing.assert_allclose(H, Hfd, rtol=1e-4, atol=1e-5 * np.abs(H).max())
    np.testing.assert_allclose(H, H.T, atol=1e-9 * np.abs(H).max())


def test_newton_direction_solves_system():
    sc = scenes.cube_stack(4, seed=1)
    p = default_params(d_hat=1e-3, mu=0.5)
    w = O.OracleWorld(sc, p)
    g, H = w.assemble(dense=True)
    rc, dq = w.newton_direction()
    assert rc == 0
    ref = -np.linalg.solve(H, g)
    np.testing.assert_allclose(dq, ref, rtol=1e-8, atol=1e-10 * np.abs(ref).max())


def test_free_fall_closed_form():
    sc = scenes.cube_stack(1, seed=1, y0=5.0)
    sc["dqdt"][0, :3] = [0.3, 1.0, -0.2]
    p = default_params(d_hat=1e-3)
    w = O.OracleWorld(sc, p)
    dt, gy = 0.01, -9.8
    for n in range(1, 21):
        rc, st = w.step()
        assert rc == 0
        q, qd = w.state()
        # implicit Euler: v_n = v_0 + n dt g, y_n = y_0 + dt sum v_k
        assert qd[0, 1] == pytest.approx(1.0 + n * dt * gy, abs=1e-9)
        assert q[0, 1] == pytest.approx(5.0 + n * dt * 1.0 + dt * dt * gy * n * (n + 1) / 2, abs=1e-9)
        assert q[0, 0] == pytest.approx(sc["q0"][0, 0] + 0.3 * n * dt, abs=1e-9)


def test_resting_cube_balances_gravity():
    d_hat, kappa, m, dt = 1e-3, 1e-4, 1000.0, 0.01
    sc = scenes.cube_stack(1, seed=1, jitter=0.0, y0=0.02)
    p = default_params(d_hat=d_hat, kappa=kappa)
    w = O.OracleWorld(sc, p)
    for _ in range(150):
        rc, st = w.step()
        assert rc == 0
    q, qd = w.state()
    gap = q[0, 1] - 0.5 + 0.5

    # 4 bottom vertices: m g dt^2 = -4 B'(gap^2) 2 gap, solved by bisection
    def f(s):
        B1 = O.barrier(s * s, d_hat, kappa)[1]
        return m * 9.8 * dt * dt + 4 * B1 * 2 * s

    lo, hi = 1e-12, math.sqrt(d_hat) * (1 - 1e-12)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) < 0:
            lo = mid
        else:
            hi = mid
    assert gap == pytest.approx(lo, rel=1e-3)
    assert abs(qd[0, 1]) < 1e-3


def test_friction_decelerates_sliding_cube():
    sc = scenes.cube_stack(1, seed=1, jitter=0.0, y0=0.02)
    # tight Newton tolerance: with the reference default (||dq|| < 1e-2, config.json:57)
    # a CCD-truncated first iteration already counts as converged and the step loses
    # most of its momentum
    p = default_params(d_hat=1e-3, mu=0.5, tol=1e-6)
    w = O.OracleWorld(sc, p)
    # settle first (no sliding)
    for _ in range(60):
        assert w.step()[0] == 0
    q, qd = w.state()
    qd[0, 0] = 1.0
    w.set_state(q, qd)
    for _ in range(10):
        assert w.step()[0] == 0
    q, qd = w.state()
    # Coulomb: dv = mu g dt per step while sliding
    assert qd[0, 0] == pytest.approx(1.0 - 0.5 * 9.8 * 0.01 * 10, rel=0.03)
