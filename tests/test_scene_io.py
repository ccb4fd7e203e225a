"""Scene JSON, config.json, OBJ meshes and trace files in the reference's
formats (adb_amd/scene_io.py), on the reference's own scene files
(test_cases/*.json and config.json, copied under tests/golden/scenes).

* Loader: body count, positions, velocities, angular velocities as
  dqdt[1..3] = skew(omega) (model/geometry.h:28-35) and rotations (theta about
  z; euler through rotation(cc, bb, aa), model/geometry.cpp:248-264) read
  straight from the JSON; an OBJ scene (cards.json) with its two meshes
  written as OBJ files (the reference's assets are not in its repository).
* OBJ reader: fan triangulation of polygons, `v/vt/vn` tokens, negative
  indices; edges as AffineObject::extract_edges (model/affine_obj.cpp:5-23).
* config.json -> adb_params field by field (view/control.cpp:20-107).
* Traces: the byte layout of player_save (model/mutils.cpp:293-329) and a
  resumed run (starting_ts, view/control.cpp:84-90) equal to the
  uninterrupted run, file for file (oracle world: any step/state/set_state
  object runs the loop).
* The friction scene steps in the oracle: the two cubes launched at 5 m/s slow
  down at mu g on the ground.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as O
from adb_amd import scene_io, scenes
from adb_amd.capi import default_params

HERE = os.path.dirname(os.path.abspath(__file__))
SC = os.path.join(HERE, "golden", "scenes")


def load(name):
    with open(os.path.join(SC, name)) as fh:
        return json.load(fh)


@pytest.mark.parametrize("name", ["10b.json", "friction.json", "static_friction.json", "ground_friction.json",
                                  "case.json", "case2.json"])
def test_cube_scene_fields(name):
    data = load(name)
    sc = scene_io.load_scene_json(os.path.join(SC, name))
    assert len(sc["mass"]) == len(data)
    for b, it in enumerate(data):
        q, qd = sc["q0"][b], sc["dqdt"][b]
        assert np.array_equal(q[:3], np.asarray(it["p"], dtype=float))
        assert np.array_equal(qd[:3], np.asarray(it["p_dot"], dtype=float))
        w = np.asarray(it["omega"], dtype=float)
        A, dA = q[3:].reshape(3, 3).T, qd[3:].reshape(3, 3).T     # columns a1, a2, a3
        # the k-th column of dA is -(e_k x w) = w x e_k: dA = [w]x I
        np.testing.assert_allclose(dA, np.cross(w, np.eye(3), axisb=0, axisc=0), atol=0)
        th = math.radians(it["theta"])
        np.testing.assert_allclose(A, [[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]],
                                   atol=1e-15)
        assert sc["mass"][b] == it.get("mass", 1000.0) and sc["Ic"][b] == sc["mass"][b] / 12
    assert sc["vert_offset"][-1] == 8 * len(data) and sc["edge_offset"][-1] == 18 * len(data)


def card_meshes(folder):
    os.makedirs(os.path.join(folder, "assets"), exist_ok=True)
    v, f, _ = scenes.cube_mesh()
    scene_io.write_obj(os.path.join(folder, "assets", "cube.obj"), v, f)
    pv = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], dtype=float) * 0.5
    with open(os.path.join(folder, "assets", "plane.obj"), "w") as fh:     # one quad, v/vt/vn tokens
        for p in pv:
            fh.write("v %g %g %g\n" % tuple(p))
        fh.write("vt 0 0\nvn 0 1 0\nf 1/1/1 2/1/1 3/1/1 4/1/1\n")


def test_obj_scene(tmp_path):
    card_meshes(str(tmp_path))
    data = load("cards.json")
    sc = scene_io.load_scene_json(data, obj_dir=str(tmp_path))
    assert len(sc["mass"]) == len(data) == 157
    nv = np.diff(sc["vert_offset"])
    for b, it in enumerate(data):
        assert nv[b] == (8 if it["obj"].endswith("cube.obj") else 4)
        if "euler" in it:
            aa, bb, cc = (math.radians(x) for x in it["euler"])
            R = scenes.euler_rotation(cc, bb, aa)
            np.testing.assert_allclose(sc["q0"][b, 3:].reshape(3, 3).T, R, atol=1e-15)
        assert sc["mass"][b] == it["mass"]
    # the quad was fanned into 2 triangles with 5 unique edges
    plane = next(b for b, it in enumerate(data) if it["obj"].endswith("plane.obj"))
    assert np.diff(sc["face_offset"])[plane] == 2 and np.diff(sc["edge_offset"])[plane] == 5


def test_obj_reader_round_trip(tmp_path):
    for v, f, e in (scenes.icosphere_mesh(2, 0.3), scenes.torus_mesh(), scenes.cube_mesh((0.1, 0.002, 0.15))):
        p = str(tmp_path / "m.obj")
        scene_io.write_obj(p, v, f)
        v2, f2, e2 = scene_io.read_obj(p)
        assert np.array_equal(v, v2) and np.array_equal(f, f2) and np.array_equal(e, e2)
    with open(p, "w") as fh:
        fh.write("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf -4 -3 -2 -1\n")
    v, f, e = scene_io.read_obj(p)
    assert f.tolist() == [[0, 1, 2], [0, 2, 3]] and len(e) == 5


def test_config_json_to_params():
    d = load("config.json")
    p = scene_io.params_from_config(os.path.join(SC, "config.json"))
    assert p.dt == d["dt"] and p.max_iter == d["max_iter"] and p.mu == d["mu"] and p.backoff == d["backoff"]
    assert list(p.gravity) == [0.0, -9.8, 0.0] and p.kappa == d["kappa"] and p.evh == d["evh"]
    assert p.enable_ground == int(d["ground"]) and p.upper_bound == int(d["uppper_bound"]) and p.iaabb == d["iaabb"]
    assert p.tol == d["params_double"]["tol"] and p.c1 == d["params_double"]["c1"]
    assert p.sh_dx == d["spatial hashing"]["dx"] and p.sh_bits == d["spatial hashing"]["xyz_bits"]
    assert p.d_hat == 1e-4 and p.kappa_ortho == 1e9     # hard-coded in the reference


def test_trace_layout(tmp_path):
    q = np.arange(24.0).reshape(2, 12)
    qd = -np.arange(24.0).reshape(2, 12)
    scene_io.write_trace(str(tmp_path), 7, q, qd)
    raw = (tmp_path / "7").read_bytes()
    assert len(raw) == 2 * 24 * 8
    a = np.frombuffer(raw, dtype="=f8").reshape(2, 24)
    assert np.array_equal(a[:, :12], q) and np.array_equal(a[:, 12:], qd)   # q0[0..3], then dqdt[0..3]
    q2, qd2 = scene_io.read_trace(str(tmp_path), 7, 2)
    assert np.array_equal(q2, q) and np.array_equal(qd2, qd)
    with pytest.raises(ValueError):
        scene_io.read_trace(str(tmp_path), 7, 3)


def test_trace_resume_equals_uninterrupted_run(tmp_path):
    sc = scene_io.load_scene_json(os.path.join(SC, "friction.json"))
    p = default_params(d_hat=1e-3, mu=0.2)
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    w = O.OracleWorld(sc, p)
    assert scene_io.run_with_trace(w, a, 10) == 10
    # a second run stopped at 5, then a fresh world resumed from trace 5
    w1 = O.OracleWorld(sc, p)
    scene_io.run_with_trace(w1, b, 5)
    w2 = O.OracleWorld(sc, p)
    assert scene_io.run_with_trace(w2, b, 10, starting_ts=5) == 10
    for ts in range(1, 11):
        assert (tmp_path / "a" / str(ts)).read_bytes() == (tmp_path / "b" / str(ts)).read_bytes(), ts


def test_friction_scene_slider_decelerates():
    # cube 1 starts at 5 m/s along x, 2 cm above the ground, alone; once the
    # barrier has settled it slides at a deceleration of mu g
    sc = scene_io.load_scene_json(os.path.join(SC, "friction.json"))
    mu, g, dt = 0.2, 9.8, 0.01
    w = O.OracleWorld(sc, default_params(d_hat=1e-3, mu=mu, tol=1e-6))
    v = []
    for _ in range(30):
        assert w.step()[0] == 0
        v.append(w.state()[1][1, 0])
    assert v[29] - v[19] == pytest.approx(-mu * g * dt * 10, rel=0.03)
    assert v[29] > 0.0
