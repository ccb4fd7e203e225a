"""Batched independent scenes (adb_world_create_batched) on the GPU.

* 6 small scenes of cubes and subdivision-2 icospheres (8 to 20 bodies),
  stepped together, against each scene stepped alone by the oracle: per-scene
  Newton counts equal and states within 1e-8 of the scene's extent over 12
  steps; and against each scene stepped alone on the GPU.
* The 256-scene workload (U(8, 200) bodies per scene, cubes or 162-vertex
  icospheres, dropped onto the ground, FP32 broad-phase boxes): 50 steps with
  every scene converging and every vertex above the ground at the end; 3
  sampled scenes against their own single-scene worlds over the first 20
  steps (Newton counts equal, states within 1e-8 of the extent).
* Bodies of different scenes never interact even when the scenes overlap in
  space: two copies of one scene at the same place step exactly like one.
"""
import numpy as np
import pytest

import oracle_lib as O
from adb_amd import scenes
from adb_amd.capi import default_params
from adb_amd.world import World

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def world_vertices(sc, q):
    vo = sc["vert_offset"]
    body = np.repeat(np.arange(len(vo) - 1), np.diff(vo))
    A = q[body, 3:].reshape(-1, 3, 3)
    return q[body, :3] + np.einsum("nki,nk->ni", A, sc["rest_verts"])


def params(fp32=0):
    return default_params(d_hat=1e-3, kappa=1e-4, mu=0.3, dt=0.01, broad_phase_fp32=fp32)


def split(full, counts):
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    return off, [scenes.subset(full, range(off[s], off[s + 1])) for s in range(len(counts))]


def test_small_batch_matches_each_scene_alone():
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    full, counts = scenes.batched_scenes(6, seed=11, nmin=8, nmax=20)
    off, parts = split(full, counts)
    p = params()
    bw = World(full, p, scene_offsets=off)
    alone = [O.OracleWorld(sc, p) for sc in parts]
    galone = [World(sc, p) for sc in parts]
    ext = [float(np.ptp(world_vertices(sc, sc["q0"]), axis=0).max()) for sc in parts]
    for step in range(12):
        rc, st = bw.step()
        newton, ls, status, contacts = bw.batch_stats()
        q, _ = bw.state()
        for s, sc in enumerate(parts):
            ro, so = alone[s].step()
            rg, sg = galone[s].step()
            assert status[s] == ro == rg == 0
            assert newton[s] == so.newton_iters == sg.newton_iters, (step, s)
            assert contacts[s] == so.n_pt + so.n_ee + so.n_vg, (step, s)
            qs = q[off[s]:off[s + 1]]
            qo = alone[s].state()[0]
            assert np.abs(world_vertices(sc, qs) - world_vertices(sc, qo)).max() <= 1e-8 * ext[s], (step, s)
            assert np.abs(world_vertices(sc, qs) - world_vertices(sc, galone[s].state()[0])).max() <= 1e-8 * ext[s]
    assert contacts.sum() > 0


def test_overlapping_scenes_do_not_interact():
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    one = scenes.cube_stack(4, seed=2)
    full = scenes.subset(one, list(range(4)) + list(range(4)))     # the same bodies twice, same place
    p = params()
    bw = World(full, p, scene_offsets=[0, 4, 8])
    sw = World(one, p)
    for step in range(10):
        bw.step()
        sw.step()
        q, _ = bw.state()
        qs, _ = sw.state()
        assert np.array_equal(q[:4], qs) and np.array_equal(q[4:], qs), step


def test_256_scenes_50_steps():
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    full, counts = scenes.batched_scenes(256, seed=3)
    off, parts = split(full, counts)
    p = params(fp32=1)
    bw = World(full, p, scene_offsets=off)
    sample = [int(np.argmin(counts)), int(np.argmax(counts)), 128]
    alone = {s: World(parts[s], p) for s in sample}
    ground = p.ground_y
    for step in range(50):
        rc, st = bw.step()
        newton, ls, status, contacts = bw.batch_stats()
        assert np.all(status == 0), step
        q, _ = bw.state()
        assert np.isfinite(q).all()
        if step >= 20:
            continue
        for s in sample:
            rs, ss = alone[s].step()
            assert newton[s] == ss.newton_iters, (step, s)
            sc = parts[s]
            ext = float(np.ptp(world_vertices(sc, sc["q0"]), axis=0).max())
            # The batched PCG runs one Krylov space over the union of the
            # scenes and stops when every scene's residual is below its
            # tolerance; a single-scene PCG stops on its own.  Iterates differ
            # at the solver tolerance (1e-12 relative), and the pile grows
            # that: ~1e-8 of the extent by step 30, then a contact that enters
            # in one run a Newton iteration before the other makes the runs
            # part.  The first 20 steps are compared; all 50 must succeed.
            dev = np.abs(world_vertices(sc, q[off[s]:off[s + 1]]) - world_vertices(sc, alone[s].state()[0])).max()
            assert dev <= 1e-8 * ext, (step, s, dev)
    for s in range(len(counts)):
        sc = parts[s]
        assert world_vertices(sc, q[off[s]:off[s + 1]])[:, 1].min() > ground
    assert contacts.sum() > 1000 and sum(counts) == bw.B
