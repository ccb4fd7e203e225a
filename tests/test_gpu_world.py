"""GPU parity of the world-level Newton contact step against the oracle.

* Contact sets (PT, EE, VG) after canonical sorting: bit-exact.
* CCD step bound: bit-exact.
* Assembled gradient / Hessian: within 1e-9 of the largest entry (FP64 sums
  in a different order, device log(), Jacobi eigensolver).
* Newton direction: PCG (rtol 1e-12) vs the oracle's sparse Cholesky, 1e-7
  relative.
* Trajectories (tight Newton tolerance so both converge): max vertex-position
  deviation over the run within 1e-8 of the scene extent.
All calls go through the C ABI of adb_amd/libadb.so.
"""
import numpy as np
import pytest

import oracle_lib as O
from adb_amd import scenes
from adb_amd.capi import default_params
from adb_amd.world import World

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _require_gpu():
    assert torch.cuda.is_available(), "GPU tests need an MI355X"


def small_spheres(n=12, seed=5):
    # subdiv-2 icospheres, tightly packed so that the d_hat band has contacts
    return scenes.sphere_pack(n, seed=seed, subdiv=2, rmin=0.2, rmax=0.3, half=0.8, ytop=1.5, gap=0.02, vsigma=0.3)


def big_pair():
    # two subdivision-3 icospheres (642 / 1280 / 1920) nearly touching, with a
    # d_hat band wide enough that more primitives per side meet the overlap
    # region than one LDS staging round holds (exercises the streamed rounds)
    v, f, e = scenes.icosphere_mesh(3, 0.1)
    rng = scenes.PCG32(9)
    bodies = [dict(verts=v, tris=f, edges=e, R=rng.rotation(), p=(0.0, 0.0, 0.0)),
              dict(verts=v, tris=f, edges=e, R=rng.rotation(), p=(0.004, 0.212, 0.003))]
    return scenes.build_scene(bodies)


SCENES = {
    "random_cubes": (lambda: scenes.random_cubes(50, seed=7, box=2.0), dict(d_hat=0.05)),
    "cube_stack": (lambda: scenes.cube_stack(10, seed=1), dict(d_hat=1e-3, mu=0.5)),
    "spheres": (lambda: small_spheres(), dict(d_hat=4e-3, mu=0.3)),
}
# contact sets and step bounds: every scene with FP64 and FP32 (outward-rounded)
# broad-phase boxes; the FP64 predicate decides in both, so both must be exact
SET_CASES = [(n, fp32) for n in list(SCENES) + ["big_pair"] for fp32 in (0, 1)]
SET_SCENES = dict(SCENES, big_pair=(big_pair, dict(d_hat=2e-2)))


def keys(a):
    return a.tolist()


@pytest.mark.parametrize("name,fp32", SET_CASES)
def test_collision_sets_bit_exact(name, fp32):
    _require_gpu()
    mk, kw = SET_SCENES[name]
    sc = mk()
    p = default_params(broad_phase_fp32=fp32, **kw)
    gw, ow = World(sc, p), O.OracleWorld(sc, p)
    for vtn, dq in [(1, None), (2, np.random.default_rng(3).normal(0, 0.02, size=(len(sc["mass"]), 12)))]:
        g = gw.collision_set(vtn, dq)
        o = ow.collision_set(vtn, dq)
        for a, b, what in zip(g, o, ("pt", "ee", "vg")):
            assert keys(a) == keys(b), (name, vtn, what, len(a), len(b))
    assert len(o[0]) + len(o[1]) > 0


@pytest.mark.parametrize("name,fp32", SET_CASES)
def test_ccd_bit_exact(name, fp32):
    _require_gpu()
    mk, kw = SET_SCENES[name]
    sc = mk()
    p = default_params(broad_phase_fp32=fp32, **kw)
    gw, ow = World(sc, p), O.OracleWorld(sc, p)
    rng = np.random.default_rng(11)
    hits = 0
    for scale in (0.01, 0.1, 0.5):
        dq = rng.normal(0, scale, size=(len(sc["mass"]), 12))
        tg, to = gw.toi(dq), ow.toi(dq)
        assert tg == to, (name, scale, tg, to)
        hits += tg < 1.0
    assert hits > 0


@pytest.mark.parametrize("name", ["cube_stack", "spheres"])
def test_assembly_matches_oracle(name):
    _require_gpu()
    mk, kw = SCENES[name]
    sc = mk()
    p = default_params(**kw)
    gw, ow = World(sc, p), O.OracleWorld(sc, p)
    gg, Hg = gw.assemble(dense=True)
    go, Ho = ow.assemble(dense=True)
    np.testing.assert_allclose(gg, go, rtol=0, atol=1e-9 * np.abs(go).max())
    np.testing.assert_allclose(Hg, Ho, rtol=0, atol=1e-9 * np.abs(Ho).max())


@pytest.mark.parametrize("name", ["cube_stack", "spheres"])
def test_newton_direction_matches_oracle(name):
    _require_gpu()
    mk, kw = SCENES[name]
    sc = mk()
    p = default_params(**kw)
    gw, ow = World(sc, p), O.OracleWorld(sc, p)
    rg, dg = gw.newton_direction()
    ro, do = ow.newton_direction()
    assert rg == 0 and ro == 0
    np.testing.assert_allclose(dg, do, rtol=0, atol=1e-7 * np.abs(do).max())


def world_vertices(sc, q):
    out = []
    for b in range(len(sc["mass"])):
        v0, v1 = sc["vert_offset"][b], sc["vert_offset"][b + 1]
        A = q[b, 3:].reshape(3, 3).T
        out.append(sc["rest_verts"][v0:v1] @ A.T + q[b, :3])
    return np.concatenate(out)


@pytest.mark.parametrize("name,steps", [("cube_stack", 20), ("spheres", 15)])
def test_trajectory_matches_oracle(name, steps):
    _require_gpu()
    mk, kw = SCENES[name]
    sc = mk()
    p = default_params(tol=1e-7, **kw)
    gw, ow = World(sc, p), O.OracleWorld(sc, p)
    X0 = world_vertices(sc, sc["q0"])
    extent = float(np.max(X0.max(0) - X0.min(0)))
    worst = 0.0
    for s in range(steps):
        rg, sg = gw.step()
        ro, so = ow.step()
        assert rg == ro == 0, (s, rg, ro)
        qg, _ = gw.state()
        qo, _ = ow.state()
        dev = np.abs(world_vertices(sc, qg) - world_vertices(sc, qo)).max()
        worst = max(worst, dev)
        assert sg.newton_iters == so.newton_iters, (s, sg.newton_iters, so.newton_iters)
    assert worst <= 1e-8 * extent, worst


def test_nan_state_reports_error():
    _require_gpu()
    sc = scenes.cube_stack(2, seed=1)
    gw = World(sc, default_params(d_hat=1e-3))
    q, qd = gw.state()
    q[0, 0] = np.nan
    gw.set_state(q, qd)
    st = gw.lib.adb_step(gw.h, None)
    assert st == -3   # ADB_E_NAN, and the process keeps running (the reference calls exit(1))


# Spatial-hash culling (iaabb = 0: contact sets and CCD through the hash grid,
# model/col_set.cpp, model/upper_bound.cpp) against the oracle's iAABB sets
# and bounds: the reference grid, a fine grid (primitives span many cells)
# and an aliasing one (2 bits: cells 0.4 apart share a slot).
HASH_GRIDS = [dict(sh_dx=0.8, sh_bits=6, sh_min_xyz=-10.0),
              dict(sh_dx=0.05, sh_bits=6, sh_min_xyz=-10.0),
              dict(sh_dx=0.1, sh_bits=2, sh_min_xyz=-10.0)]


@pytest.mark.parametrize("grid", range(len(HASH_GRIDS)))
@pytest.mark.parametrize("name", list(SET_SCENES))
def test_hash_sets_and_ccd_bit_exact(name, grid):
    _require_gpu()
    mk, kw = SET_SCENES[name]
    sc = mk()
    gw = World(sc, default_params(iaabb=0, **HASH_GRIDS[grid], **kw))
    ow = O.OracleWorld(sc, default_params(**kw))
    for vtn, dq in [(1, None), (2, np.random.default_rng(3).normal(0, 0.02, size=(len(sc["mass"]), 12)))]:
        g = gw.collision_set(vtn, dq)
        o = ow.collision_set(vtn, dq, method=1)
        for a, b, what in zip(g, o, ("pt", "ee", "vg")):
            assert keys(a) == keys(b), (name, grid, vtn, what, len(a), len(b))
    assert len(o[0]) + len(o[1]) > 0
    rng = np.random.default_rng(11)
    hits = 0
    for scale in (0.01, 0.1, 0.5):
        dq = rng.normal(0, scale, size=(len(sc["mass"]), 12))
        tg, to = gw.toi(dq), ow.toi(dq, method=1)
        assert tg == to, (name, grid, scale, tg, to)
        hits += tg < 1.0
    assert hits > 0


def test_hash_mode_trajectory_equals_iaabb_mode():
    # the same pairs reach the same predicates, so the steps are bit-identical
    _require_gpu()
    mk, kw = SCENES["spheres"]
    sc = mk()
    qs = []
    for mode in (3, 0):
        gw = World(sc, default_params(iaabb=mode, sh_dx=0.1, sh_bits=5, **kw))
        for s in range(8):
            rc, st = gw.step()
            assert rc == 0, (mode, s, rc)
        qs.append(gw.state()[0])
    assert np.array_equal(qs[0], qs[1])


@pytest.mark.parametrize("name,steps", [("cube_stack", 8), ("spheres", 6)])
def test_general_paths_give_identical_steps(name, steps, monkeypatch):
    # the paths taken above a size (contact-set keys radix-sorted instead of
    # ranked, the assembly's pair count read by the host instead of bounded,
    # its block rows and chunk table from the multi-launch kernels instead of
    # one workgroup, the body boxes from one workgroup per body instead of one
    # wavefront) give the same steps bit for bit
    _require_gpu()
    mk, kw = SCENES[name]
    sc = mk()
    runs = []
    for general in ("0", "1"):
        monkeypatch.setenv("ADB_GENERAL_PATHS", general)
        gw = World(sc, default_params(**kw))
        counts = []
        for s in range(steps):
            rc, st = gw.step()
            assert rc == 0, (general, s, rc)
            counts.append((st.newton_iters, st.ls_trials, st.pcg_iters, st.n_pt, st.n_ee, st.n_vg))
        runs.append((gw.state(), counts))
    assert runs[0][1] == runs[1][1]
    assert sum(c[3] + c[4] for c in runs[0][1]) > 0
    assert np.array_equal(runs[0][0][0], runs[1][0][0]) and np.array_equal(runs[0][0][1], runs[1][0][1])


@pytest.mark.parametrize("name,steps", [("cube_stack", 10), ("spheres", 8)])
def test_line_search_candidate_sets_equal_full_passes(name, steps, monkeypatch):
    # trial sets filtered from the candidates of the search segment equal the
    # sets of a full culling pass at each trial point: identical steps
    _require_gpu()
    mk, kw = SCENES[name]
    sc = mk()
    runs = []
    for cand in ("1", "0"):
        monkeypatch.setenv("ADB_LS_CANDIDATES", cand)
        gw = World(sc, default_params(**kw))
        counts = []
        for s in range(steps):
            rc, st = gw.step()
            assert rc == 0, (cand, s, rc)
            counts.append((st.newton_iters, st.ls_trials, st.pcg_iters, st.n_pt, st.n_ee, st.n_vg))
        runs.append((gw.state()[0], counts))
    assert runs[0][1] == runs[1][1]
    assert sum(c[1] for c in runs[0][1]) > 0 and sum(c[3] + c[4] for c in runs[0][1]) > 0
    assert np.array_equal(runs[0][0], runs[1][0])
