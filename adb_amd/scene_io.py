"""Scene, configuration, mesh and trace files in the reference's formats.

* Scene JSON (test_cases/*.json, read by customize(),
  test_cases/spinning_cube.cpp:56-127): a list of bodies with p, p_dot, omega,
  theta (degrees about z) or euler (degrees), optional mass (Ic = mass / 12)
  and optional obj (an OBJ mesh; a unit cube otherwise).
* config.json (read by reset(), view/control.cpp:20-107) -> adb_params.
* OBJ meshes (AffineObject, model/affine_obj.cpp:5-23, loaded through assimp
  at view/model.h:50-60): `v` and `f` records, polygons fanned into
  triangles, edges = the unique triangle edges.
* Trace files (player_save / player_load, model/mutils.cpp:257-329): one file
  per timestep, `<folder>/<ts>`, holding per body q0 (p, a1, a2, a3: 12
  doubles) then dqdt (12 doubles), native-endian float64.  The viewer saves
  file ts after the step that advances globals.ts to ts
  (view/model_view.cpp:283-289, model/abd.cpp:116) and resumes from
  starting_ts by loading that file (view/control.cpp:84-90).
"""
import json
import math
import os

import numpy as np

from . import scenes
from .capi import default_params


# --------------------------------------------------------------------------- OBJ
def read_obj(path):
    """(verts V x 3, tris F x 3 int32, edges E x 2 int32) of an OBJ file.
    Vertices are shared as indexed in the file (assimp without
    aiProcess_JoinIdenticalVertices splits them per face; the contact
    primitives of an indexed mesh are the same surface)."""
    verts, tris = [], []
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "v":
                verts.append([float(x) for x in parts[1:4]])
            elif parts[0] == "f":
                idx = []
                for tok in parts[1:]:
                    i = int(tok.split("/")[0])
                    idx.append(i - 1 if i > 0 else len(verts) + i)
                for k in range(1, len(idx) - 1):
                    tris.append((idx[0], idx[k], idx[k + 1]))
    v = np.array(verts, dtype=np.float64).reshape(-1, 3)
    f = np.array(tris, dtype=np.int32).reshape(-1, 3)
    if len(f) and (f.min() < 0 or f.max() >= len(v)):
        raise ValueError("%s: face index out of range" % path)
    return v, f, scenes.unique_edges(f)


def write_obj(path, verts, tris):
    with open(path, "w") as fh:
        for p in np.asarray(verts, dtype=np.float64).reshape(-1, 3):
            fh.write("v %.17g %.17g %.17g\n" % tuple(p))
        for t in np.asarray(tris).reshape(-1, 3):
            fh.write("f %d %d %d\n" % (t[0] + 1, t[1] + 1, t[2] + 1))


# --------------------------------------------------------------------------- scene JSON
def _rotation(it):
    if "euler" in it:
        a, b, c = (x / 180.0 * math.pi for x in it["euler"])
        # rotation(cc, bb, aa) of model/geometry.cpp:248-264 (test_cases/spinning_cube.cpp:105-111)
        return scenes.euler_rotation(c, b, a)
    th = it["theta"] / 180.0 * math.pi
    s, c = math.sin(th), math.cos(th)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def load_scene_json(path_or_list, obj_dir=None, meshes=None):
    """A scene dict (adb_scene layout) from the reference's scene JSON.

    OBJ bodies are read from `meshes` ({name: (verts, tris[, edges])}) when
    given, else from obj_dir joined with the entry's path.  Modal bodies
    (npy / lambda entries) are not part of the affine contact step."""
    if isinstance(path_or_list, str):
        with open(path_or_list) as fh:
            data = json.load(fh)
        if obj_dir is None:
            obj_dir = os.path.dirname(os.path.abspath(path_or_list))
    else:
        data = path_or_list
    cube = scenes.cube_mesh()
    cache = {}
    bodies, spins = [], []
    for it in data:
        if "npy" in it:
            raise ValueError("modal bodies (npy / lambda) are not supported")
        if "obj" in it:
            name = it["obj"]
            if name not in cache:
                if meshes is not None and name in meshes:
                    m = meshes[name]
                    cache[name] = (np.asarray(m[0], dtype=np.float64), np.asarray(m[1], dtype=np.int32),
                                   m[2] if len(m) > 2 else scenes.unique_edges(np.asarray(m[1])))
                else:
                    cache[name] = read_obj(os.path.join(obj_dir or ".", name))
            v, f, e = cache[name]
        else:
            v, f, e = cube
        m = float(it.get("mass", 1000.0))
        wx, wy, wz = it["omega"]
        # dqdt[1..3] = skew(omega) (model/geometry.h:28-35)
        spins.append(np.concatenate([-np.array([0.0, -wz, wy]), -np.array([wz, 0.0, -wx]), -np.array([-wy, wx, 0.0])]))
        bodies.append(dict(verts=v, tris=f, edges=e, R=_rotation(it), p=it["p"], v=it["p_dot"], mass=m, Ic=m / 12.0))
    sc = scenes.build_scene(bodies)
    for b, s in enumerate(spins):
        sc["dqdt"][b, 3:] = s
    return sc


def params_from_config(path_or_dict, **overrides):
    """adb_params from the reference's config.json (view/control.cpp:20-107).
    d_hat keeps the hard-coded barrier::d_hat = 1e-4 (model/barrier.cpp:7)
    unless overridden; ortho kappa is the hard-coded 1e9."""
    if isinstance(path_or_dict, str):
        with open(path_or_dict) as fh:
            d = json.load(fh)
    else:
        d = path_or_dict
    pd = d.get("params_double", {})
    sh = d.get("spatial hashing", {})
    kw = dict(
        dt=d["dt"], max_iter=d["max_iter"], gravity=(0.0, -9.8 if d["gravity"] else 0.0, 0.0),
        line_search=int(d["line_search"]), enable_ee=int(d["ee"]), enable_pt=int(d["pt"]),
        enable_ground=int(d["ground"]), upper_bound=int(d["uppper_bound"]), mu=d["mu"], backoff=d["backoff"],
        eps_x=d["eps_x"], psd=int(d["psd"]), iaabb=d["iaabb"], evh=d["evh"], vg_fric=int(d["vg_friction"]),
        pt_fric=int(d["pt_friction"]), ee_fric=int(d["ee_friction"]), kappa=d["kappa"],
        tol=pd.get("tol", 1e-2), c1=pd.get("c1", 1e-4), max_uk=pd.get("max_uk", 0.0))
    if sh:
        kw.update(sh_min_xyz=sh["min_xyz"], sh_dx=sh["dx"], sh_bits=sh["xyz_bits"])
    kw.update(overrides)
    return default_params(**kw)


# --------------------------------------------------------------------------- traces
def trace_path(folder, ts):
    return os.path.join(folder, str(int(ts)))


def write_trace(folder, ts, q, qdot):
    """player_save: per body q0 (12) then dqdt (12), float64."""
    q = np.asarray(q, dtype=np.float64).reshape(-1, 12)
    qd = np.asarray(qdot, dtype=np.float64).reshape(-1, 12)
    np.ascontiguousarray(np.concatenate([q, qd], axis=1)).astype("=f8").tofile(trace_path(folder, ts))


def read_trace(folder, ts, n_bodies):
    """player_load: (q0, dqdt), B x 12 each."""
    a = np.fromfile(trace_path(folder, ts), dtype="=f8")
    if a.size != 24 * n_bodies:
        raise ValueError("trace %s holds %d doubles, expected %d" % (trace_path(folder, ts), a.size, 24 * n_bodies))
    a = a.reshape(n_bodies, 24)
    return a[:, :12].copy(), a[:, 12:].copy()


def run_with_trace(world, folder, ending_ts, starting_ts=-1, on_step=None):
    """The viewer's loop (view/model_view.cpp:283-289, view/control.cpp:84-90):
    resume from `<folder>/<starting_ts>` when starting_ts >= 0 (else start at
    ts = 0 and recreate the folder), then step and save `<folder>/<ts>` after
    each step until ts reaches ending_ts.  `world` has step(), state() and
    set_state(q, qdot).  Returns the last ts."""
    ts = 0
    if starting_ts >= 0:
        q, qd = read_trace(folder, starting_ts, world.B)
        world.set_state(q, qd)
        ts = starting_ts
    else:
        os.makedirs(folder, exist_ok=True)
        for name in os.listdir(folder):
            if name.isdigit():
                os.remove(os.path.join(folder, name))
    while ts < ending_ts:
        rc, st = world.step()
        ts += 1
        q, qd = world.state()
        write_trace(folder, ts, q, qd)
        if on_step:
            on_step(ts, rc, st)
    return ts
