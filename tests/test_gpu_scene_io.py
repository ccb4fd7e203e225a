"""The reference's scene files run on the GPU (through the C ABI) against the
oracle, and the trace resume (starting_ts) on the GPU.

* test_cases/10b.json, friction.json, static_friction.json (copied under
  tests/golden/scenes): 30 steps in lockstep with the oracle (equal Newton
  counts, states within 1e-8 of the scene extent, sets bit-exact every 10
  steps).
* A run resumed from its trace file 5 writes trace files 6..10 identical,
  byte for byte, to the uninterrupted run's (the device path is deterministic).
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from adb_amd import scene_io
from adb_amd.capi import default_params
from adb_amd.world import World

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
SC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")


def world_vertices(sc, q):
    vo = sc["vert_offset"]
    body = np.repeat(np.arange(len(vo) - 1), np.diff(vo))
    A = q[body, 3:].reshape(-1, 3, 3)
    return q[body, :3] + np.einsum("nki,nk->ni", A, sc["rest_verts"])


@pytest.mark.parametrize("name,mu", [("10b.json", 0.0), ("friction.json", 0.2), ("static_friction.json", 0.5)])
def test_reference_scene_lockstep(name, mu):
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    sc = scene_io.load_scene_json(os.path.join(SC, name))
    # a tight Newton tolerance: with sliding friction the reference's 1e-2
    # leaves iterates whose rounding-level differences grow past 1e-8
    p = scene_io.params_from_config(os.path.join(SC, "config.json"), d_hat=1e-3, mu=mu, enable_ground=1, tol=1e-7)
    gw, ow = World(sc, p), O.OracleWorld(sc, p)
    X = world_vertices(sc, sc["q0"])
    ext = float(np.max(X.max(0) - X.min(0)))
    for s in range(30):
        rg, sg = gw.step()
        ro, so = ow.step()
        assert rg == ro == 0 and sg.newton_iters == so.newton_iters, (name, s)
        qg, qdg = gw.state()
        qo, qdo = ow.state()
        assert np.abs(world_vertices(sc, qg) - world_vertices(sc, qo)).max() <= 1e-8 * ext, (name, s)
        if (s + 1) % 10 == 0:
            ow.set_state(qg, qdg)
            for a, b in zip(gw.collision_set(), ow.collision_set()):
                assert a.tolist() == b.tolist(), (name, s)
            ow.set_state(qo, qdo)


def test_trace_resume_on_gpu(tmp_path):
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    sc = scene_io.load_scene_json(os.path.join(SC, "friction.json"))
    p = default_params(d_hat=1e-3, mu=0.2)
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    scene_io.run_with_trace(World(sc, p), a, 10)
    scene_io.run_with_trace(World(sc, p), b, 5)
    scene_io.run_with_trace(World(sc, p), b, 10, starting_ts=5)
    for ts in range(1, 11):
        assert (tmp_path / "a" / str(ts)).read_bytes() == (tmp_path / "b" / str(ts)).read_bytes(), ts
