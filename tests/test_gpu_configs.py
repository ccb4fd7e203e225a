"""GPU parity on the benchmark configurations, through the C ABI.

* 10 stacked unit cubes (mu 0.5, d_hat 1e-3, the reference's Newton tolerance
  1e-2): all 100 steps against the oracle, equal Newton counts, states within
  1e-8 of the scene extent.
* 2000-icosphere workload, downsized: 48 subdivision-3 spheres with the same
  radius law dropped into a 0.9 x 0.9 box so they pile within 12 steps;
  per-step Newton counts, states and end-of-step contact sets against the
  oracle, in FP64 and with FP32 broad-phase boxes.
* 2000-icosphere workload at full size: 25 steps (every step ADB_OK; the
  FP32-box and FP64 runs bit-identical); at the reached state the contact sets
  and the CCD bound of the Newton direction are bit-exact against the oracle,
  and the sets are sorted and free of duplicates.
* The same workload at a late, piled-up state (step 70, committed fixture):
  full-size sets bit-exact in both precisions; on the 60 bodies nearest the
  pile's centre the assembled gradient and Hessian (1e-9 of the largest
  entry), the Newton direction (1e-7) and one whole step against the oracle.
* Torus chain (differentiable workload's forward pass): 8 links against the
  oracle over 20 steps; the 64-link chain for 100 steps (pinned link fixed).
* Leaning card lattice: 36 cards against the oracle over 20 steps; the
  10k-card lattice for 20 steps with sets bit-exact at the end.
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from adb_amd import scenes
from adb_amd.capi import default_params
from adb_amd.world import World

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
HERE = os.path.dirname(os.path.abspath(__file__))
PILE = os.path.join(HERE, "golden", "c2_pile_step70.npz")


def _require_gpu():
    assert torch.cuda.is_available(), "GPU tests need an MI355X"


def world_vertices(sc, q):
    vo = sc["vert_offset"]
    body = np.repeat(np.arange(len(vo) - 1), np.diff(vo))
    A = q[body, 3:].reshape(-1, 3, 3)           # rows a1, a2, a3
    return q[body, :3] + np.einsum("nki,nk->ni", A, sc["rest_verts"])


def extent(sc):
    X = world_vertices(sc, sc["q0"])
    return float(np.max(X.max(0) - X.min(0)))


def assert_sorted_unique(a, what):
    if len(a) > 1:
        t = [tuple(r) for r in a.tolist()]
        assert all(x < y for x, y in zip(t, t[1:])), what + " not strictly sorted"


def lockstep(sc, p, steps, sets_every=1, tag=""):
    """GPU and oracle step by step: equal status and Newton counts, state
    deviation, and (every `sets_every` steps) equal end-of-step sets."""
    gw, ow = World(sc, p), O.OracleWorld(sc, p)
    ext = extent(sc)
    worst = 0.0
    contacts = 0
    for s in range(steps):
        rg, sg = gw.step()
        ro, so = ow.step()
        assert rg == ro == 0, (tag, s, rg, ro)
        assert sg.newton_iters == so.newton_iters, (tag, s, sg.newton_iters, so.newton_iters)
        assert (sg.n_pt, sg.n_ee, sg.n_vg) == (so.n_pt, so.n_ee, so.n_vg), (tag, s)
        qg, _ = gw.state()
        qo, _ = ow.state()
        dev = float(np.abs(world_vertices(sc, qg) - world_vertices(sc, qo)).max())
        worst = max(worst, dev)
        assert dev <= 1e-8 * ext, (tag, s, dev, ext)
        contacts = max(contacts, sg.n_pt + sg.n_ee + sg.n_vg)
        if sets_every and (s + 1) % sets_every == 0:
            qdo = ow.state()[1]
            ow.set_state(qg, gw.state()[1])   # the oracle's sets at the GPU state
            for a, b, what in zip(gw.collision_set(), ow.collision_set(), ("pt", "ee", "vg")):
                assert a.tolist() == b.tolist(), (tag, s, what, len(a), len(b))
            ow.set_state(qo, qdo)
    return worst, contacts


def test_c1_ten_cubes_hundred_steps():
    _require_gpu()
    sc = scenes.cube_stack(10, seed=1)
    p = default_params(d_hat=1e-3, mu=0.5, dt=0.01)   # Newton tol 1e-2 as config.json:57
    worst, contacts = lockstep(sc, p, 100, sets_every=10, tag="C1")
    assert contacts > 0


def c2_small():
    # the 2000-icosphere radius law and mesh in a smaller box, dropped low
    return scenes.sphere_pack(48, seed=2, subdiv=3, half=0.45, ytop=0.8)


@pytest.mark.parametrize("fp32", [0, 1])
def test_c2_downsized_lockstep(fp32):
    _require_gpu()
    sc = c2_small()
    p = default_params(d_hat=1e-3, kappa=1e-4, mu=0.3, dt=0.01, broad_phase_fp32=fp32)
    worst, contacts = lockstep(sc, p, 12, sets_every=3, tag="C2-48")
    assert contacts > 5000


def test_c2_full_size_first_steps():
    _require_gpu()
    import bench
    sc = bench.workload(2000)
    states = []
    for prec in ("mixed", "fp64"):
        w = World(sc, bench.params(prec))
        for s in range(25):
            rc, st = w.step()       # raises on any status but ADB_OK
            assert st.pcg_not_converged == 0 and st.pcg_breakdown == 0, (prec, s)
        states.append(w.state())
        if prec == "mixed":
            gw = w
    assert np.array_equal(states[0][0], states[1][0]) and np.array_equal(states[0][1], states[1][1])
    ow = O.OracleWorld(sc, bench.params("fp64"))
    ow.set_state(*states[0])
    g = gw.collision_set()
    o = ow.collision_set()
    for a, b, what in zip(g, o, ("pt", "ee", "vg")):
        assert_sorted_unique(a, what)
        assert a.tolist() == b.tolist(), (what, len(a), len(b))
    assert len(g[0]) + len(g[1]) > 50000
    rc, dq = gw.newton_direction()
    assert rc == 0
    tg, to = gw.toi(dq), ow.toi(dq)
    assert tg == to, (tg, to)


def pile_state():
    if not os.path.exists(PILE):
        pytest.skip("late-pile fixture missing (tests/golden/make_c2_pile_state.py writes it)")
    z = np.load(PILE)
    return z["q"], z["qdot"]


@pytest.mark.parametrize("fp32", [0, 1])
def test_c2_pile_sets_and_ccd_full_size(fp32):
    _require_gpu()
    import bench
    q, qd = pile_state()
    sc = bench.workload(2000)
    sc["q0"], sc["dqdt"] = q, qd
    p = default_params(d_hat=1e-3, kappa=1e-4, mu=0.3, dt=0.01, broad_phase_fp32=fp32)
    gw, ow = World(sc, p), O.OracleWorld(sc, p)
    g, o = gw.collision_set(), ow.collision_set()
    for a, b, what in zip(g, o, ("pt", "ee", "vg")):
        assert_sorted_unique(a, what)
        assert a.tolist() == b.tolist(), (what, len(a), len(b))
    assert len(g[0]) + len(g[1]) > 1_000_000
    rng = np.random.default_rng(70)
    dq = rng.normal(0, 2e-3, size=(2000, 12))
    dq[:, 3:] *= 0.1
    tg, to = gw.toi(dq), ow.toi(dq)
    assert tg == to and tg < 1.0, (tg, to)


def pile_crop(n=60):
    import bench
    q, qd = pile_state()
    full = bench.workload(2000)
    centre = np.array([0.0, -0.4, 0.0])
    near = np.argsort(np.linalg.norm(q[:, :3] - centre, axis=1))[:n]
    return scenes.subset(full, np.sort(near), q, qd)


def test_c2_pile_crop_assembly_direction_and_step():
    _require_gpu()
    sc = pile_crop()
    p = default_params(d_hat=1e-3, kappa=1e-4, mu=0.3, dt=0.01)
    gw, ow = World(sc, p), O.OracleWorld(sc, p)
    pt, ee, vg = ow.collision_set()
    # the dense regime: many contacts per body pair, 4-point contacts present
    X = world_vertices(sc, sc["q0"])
    types = [O.pt_distance(np.concatenate([X[v] for v in ids]))[0] for ids in _pt_ids(sc, pt[:2000])]
    assert len(pt) + len(ee) > 20000 and 6 in types
    gg, Hg = gw.assemble(dense=True)
    go, Ho = ow.assemble(dense=True)
    np.testing.assert_allclose(gg, go, rtol=0, atol=1e-9 * np.abs(go).max())
    np.testing.assert_allclose(Hg, Ho, rtol=0, atol=1e-9 * np.abs(Ho).max())
    rg, dg = gw.newton_direction()
    ro, do = ow.newton_direction()
    assert rg == 0 and ro == 0
    np.testing.assert_allclose(dg, do, rtol=0, atol=1e-7 * np.abs(do).max())
    rg, sg = gw.step()
    ro, so = ow.step()
    assert rg == ro == 0 and sg.newton_iters == so.newton_iters
    dev = float(np.abs(world_vertices(sc, gw.state()[0]) - world_vertices(sc, ow.state()[0])).max())
    assert dev <= 1e-8 * extent(sc), dev


def _pt_ids(sc, pt):
    vo, fo = sc["vert_offset"], sc["face_offset"]
    for I, v, J, f in pt:
        yield [vo[I] + v] + [vo[J] + k for k in sc["faces"][fo[J] + f]]


def chain_params():
    # barrier band (sqrt(d_hat) = 3.2 mm) thinner than the tube radius (25 mm)
    return default_params(d_hat=1e-5, kappa=1e-4, mu=0.0, dt=0.01, enable_ground=0)


def test_c4_torus_chain_downsized_lockstep():
    _require_gpu()
    sc = scenes.torus_chain(8, seed=4)
    worst, contacts = lockstep(sc, chain_params(), 20, sets_every=5, tag="C4-8")
    assert contacts > 1000


def test_c4_torus_chain_full_forward():
    _require_gpu()
    sc = scenes.torus_chain(64, seed=4)
    w = World(sc, chain_params())
    q0 = sc["q0"].copy()
    for s in range(100):
        rc, st = w.step()
    q, qd = w.state()
    assert np.array_equal(q[0], q0[0])            # the pinned link (mass < 0) never moves
    assert q[63, 1] < q0[63, 1]                  # the chain sags under gravity
    assert np.all(np.isfinite(q))
    pt, ee, vg = w.collision_set()
    assert len(pt) + len(ee) > 0 and len(vg) == 0


def test_c5_cards_downsized_lockstep():
    _require_gpu()
    sc = scenes.card_lattice(36, seed=5)
    p = default_params(d_hat=1e-3, kappa=1e-4, mu=0.6, dt=0.01, broad_phase_fp32=1)
    worst, contacts = lockstep(sc, p, 20, sets_every=5, tag="C5-36")
    assert contacts > 1000


def test_c5_cards_full_size():
    _require_gpu()
    sc = scenes.card_lattice(10000, seed=5)
    p = default_params(d_hat=1e-3, kappa=1e-4, mu=0.6, dt=0.01, broad_phase_fp32=1)
    w = World(sc, p)
    for s in range(20):
        rc, st = w.step()
    assert st.n_vg > 10000
    ow = O.OracleWorld(sc, p)
    ow.set_state(*w.state())
    for a, b, what in zip(w.collision_set(), ow.collision_set(), ("pt", "ee", "vg")):
        assert_sorted_unique(a, what)
        assert a.tolist() == b.tolist(), (what, len(a), len(b))
