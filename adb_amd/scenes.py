"""Seeded scene generators (PCG32) for the ADB contact step.

A scene is a dict of flat arrays in the layout of `adb_scene`
(include/adb_capi.h): per-body vertex/face/edge offsets, rest vertices,
LOCAL face and edge indices, mass, Ic, q0 = [p, a1, a2, a3] and dqdt.

Body conventions follow the reference: mass 1000 and Ic = mass / 12
(model/cube.h:38, model/affine_obj.h:10); cube faces from model/cube.h:15-36
split into triangles as Cube::gen_indices (model/cube.h:42-63); edges are the
unique triangle edges as AffineObject::extract_edges (model/affine_obj.cpp:5-23);
a static body has mass < 0 (model/mutils.cpp:55,71).
"""
import math

import numpy as np

MASK64 = (1 << 64) - 1


class PCG32:
    """PCG-XSH-RR 32-bit generator (O'Neill 2014)."""

    def __init__(self, seed, seq=54):
        self.state = 0
        self.inc = ((seq << 1) | 1) & MASK64
        self.next_u32()
        self.state = (self.state + seed) & MASK64
        self.next_u32()

    def next_u32(self):
        old = self.state
        self.state = (old * 6364136223846793005 + self.inc) & MASK64
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    def uniform(self, lo=0.0, hi=1.0):
        """53-bit uniform double in [lo, hi)."""
        a = self.next_u32() >> 5
        b = self.next_u32() >> 6
        u = (a * 67108864.0 + b) / 9007199254740992.0
        return lo + (hi - lo) * u

    def normal(self, mu=0.0, sigma=1.0):
        u1 = self.uniform()
        u2 = self.uniform()
        u1 = max(u1, 1e-300)
        return mu + sigma * math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def rotation(self):
        """Uniform random rotation (Shoemake's subgroup algorithm)."""
        u1, u2, u3 = self.uniform(), self.uniform(), self.uniform()
        a, b = math.sqrt(1 - u1), math.sqrt(u1)
        qx, qy = a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2)
        qz, qw = b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3)
        return quat_to_mat(qw, qx, qy, qz)


def quat_to_mat(w, x, y, z):
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def euler_rotation(a, b, c):
    """rotation(a, b, c) of model/geometry.cpp:248-264."""
    s1, s2, s3 = math.sin(a), math.sin(b), math.sin(c)
    c1, c2, c3 = math.cos(a), math.cos(b), math.cos(c)
    return np.array([
        [c1 * c2, c1 * s2 * s3 - c3 * s1, s1 * s3 + c1 * c3 * s2],
        [c2 * s1, c1 * c3 + s1 * s2 * s3, c3 * s1 * s2 - c1 * s3],
        [-s2, c2 * s3, c2 * c3],
    ])


# --------------------------------------------------------------------------- meshes
def unique_edges(tris):
    s = set()
    for t in tris:
        a, b, c = int(t[0]), int(t[1]), int(t[2])
        for u, v in ((a, b), (b, c), (a, c)):
            s.add((min(u, v), max(u, v)))
    return np.array(sorted(s), dtype=np.int32).reshape(-1, 2)


def cube_mesh(size=(1.0, 1.0, 1.0)):
    v = np.array([[-0.5, -0.5, -0.5], [-0.5, -0.5, 0.5], [-0.5, 0.5, -0.5], [-0.5, 0.5, 0.5],
                  [0.5, -0.5, -0.5], [0.5, -0.5, 0.5], [0.5, 0.5, -0.5], [0.5, 0.5, 0.5]])
    v = v * np.asarray(size)[None, :]
    quads = [(0, 1, 3, 2), (4, 5, 1, 0), (4, 0, 2, 6), (2, 3, 7, 6), (1, 5, 7, 3), (5, 4, 6, 7)]
    tris = []
    for f in quads:
        tris.append((f[0], f[1], f[2]))
        tris.append((f[2], f[3], f[0]))
    tris = np.array(tris, dtype=np.int32)
    return v, tris, unique_edges(tris)


def icosphere_mesh(subdiv, radius=1.0):
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    v = np.array(verts) * radius
    tris = np.array(faces, dtype=np.int32)
    return v, tris, unique_edges(tris)


def torus_mesh(R=0.1, r=0.025, nu=24, nv=12):
    verts = []
    for i in range(nu):
        u = 2 * math.pi * i / nu
        for j in range(nv):
            w = 2 * math.pi * j / nv
            verts.append(((R + r * math.cos(w)) * math.cos(u), (R + r * math.cos(w)) * math.sin(u), r * math.sin(w)))
    tris = []
    for i in range(nu):
        for j in range(nv):
            a = i * nv + j
            b = ((i + 1) % nu) * nv + j
            c = ((i + 1) % nu) * nv + (j + 1) % nv
            d = i * nv + (j + 1) % nv
            tris.append((a, b, c))
            tris.append((a, c, d))
    tris = np.array(tris, dtype=np.int32)
    return np.array(verts), tris, unique_edges(tris)


# --------------------------------------------------------------------------- assembly
def build_scene(bodies):
    """bodies: list of dicts with keys verts, tris, edges, R (3x3), p (3), v (3),
    omega (3, optional), mass (optional, <0 static), Ic (optional)."""
    vo, fo, eo = [0], [0], [0]
    rv, fs, es, mass, Ic, q0, dq = [], [], [], [], [], [], []
    for b in bodies:
        v, f, e = b["verts"], b["tris"], b["edges"]
        vo.append(vo[-1] + len(v))
        fo.append(fo[-1] + len(f))
        eo.append(eo[-1] + len(e))
        rv.append(np.asarray(v, dtype=np.float64))
        fs.append(np.asarray(f, dtype=np.int32))
        es.append(np.asarray(e, dtype=np.int32))
        m = b.get("mass", 1000.0)
        mass.append(m)
        Ic.append(b.get("Ic", abs(m) / 12.0))
        R = np.asarray(b.get("R", np.eye(3)), dtype=np.float64)
        q = np.concatenate([np.asarray(b["p"], dtype=np.float64), R[:, 0], R[:, 1], R[:, 2]])
        q0.append(q)
        w = np.asarray(b.get("omega", (0.0, 0.0, 0.0)), dtype=np.float64)
        # dA/dt = [w]x A, columns a_k' = w x a_k
        dA = [np.cross(w, R[:, k]) for k in range(3)]
        dq.append(np.concatenate([np.asarray(b.get("v", (0.0, 0.0, 0.0)), dtype=np.float64)] + dA))
    return {
        "vert_offset": np.array(vo, dtype=np.int32), "face_offset": np.array(fo, dtype=np.int32),
        "edge_offset": np.array(eo, dtype=np.int32),
        "rest_verts": np.concatenate(rv).reshape(-1, 3), "faces": np.concatenate(fs).reshape(-1, 3),
        "edges": np.concatenate(es).reshape(-1, 2), "mass": np.array(mass), "Ic": np.array(Ic),
        "q0": np.array(q0).reshape(-1, 12), "dqdt": np.array(dq).reshape(-1, 12),
    }


def subset(scene, bodies, q=None, qdot=None):
    """The scene restricted to the listed bodies (in that order), with their
    rest meshes, mass and Ic; q / qdot (B x 12, optional) replace q0 / dqdt."""
    bodies = [int(b) for b in bodies]
    vo, fo, eo = scene["vert_offset"], scene["face_offset"], scene["edge_offset"]
    q = scene["q0"] if q is None else np.asarray(q).reshape(-1, 12)
    qdot = scene["dqdt"] if qdot is None else np.asarray(qdot).reshape(-1, 12)
    out = []
    for b in bodies:
        out.append(dict(verts=scene["rest_verts"][vo[b]:vo[b + 1]], tris=scene["faces"][fo[b]:fo[b + 1]],
                        edges=scene["edges"][eo[b]:eo[b + 1]], p=q[b, :3], mass=float(scene["mass"][b]),
                        Ic=float(scene["Ic"][b])))
    sc = build_scene(out)
    sc["q0"] = np.array(q[bodies], dtype=np.float64).reshape(-1, 12)
    sc["dqdt"] = np.array(qdot[bodies], dtype=np.float64).reshape(-1, 12)
    return sc


def scene_counts(scene):
    return {"bodies": len(scene["mass"]), "V": int(scene["vert_offset"][-1]),
            "F": int(scene["face_offset"][-1]), "E": int(scene["edge_offset"][-1])}


# --------------------------------------------------------------------------- workloads
def cube_stack(n=10, seed=1, jitter=0.05, spacing=1.01, y0=0.01, rotate=False):
    """Stack of unit cubes above the ground (spacing as test_cases/10b.json)."""
    rng = PCG32(seed)
    v, f, e = cube_mesh()
    bodies = []
    for i in range(n):
        p = (rng.uniform(-jitter, jitter), y0 + spacing * i, rng.uniform(-jitter, jitter))
        R = rng.rotation() if rotate else np.eye(3)
        bodies.append(dict(verts=v, tris=f, edges=e, R=R, p=p))
    return build_scene(bodies)


def random_cubes(n=50, seed=7, box=3.0, dq_scale=0.0):
    """Randomly placed and rotated unit cubes in [-box, box]^3 (the fixture of
    iAABB/test.cpp:89-169, with a portable generator)."""
    rng = PCG32(seed)
    v, f, e = cube_mesh()
    bodies = []
    for i in range(n):
        R = euler_rotation(rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi), rng.uniform(0, 2 * math.pi))
        p = (rng.uniform(-box, box), rng.uniform(-box, box), rng.uniform(-box, box))
        bodies.append(dict(verts=v, tris=f, edges=e, R=R, p=p))
    return build_scene(bodies)


def sphere_pack(n=2000, seed=2, subdiv=3, rmin=0.05, rmax=0.15, half=2.0, ytop=7.5, ground_y=-0.5,
                gap=2.0 * math.sqrt(1e-3), vsigma=0.5, max_tries=200000):
    """n icospheres with r ~ U(rmin, rmax) rejection-packed in x,z in [-half, half],
    y in [ground + r + gap, ytop], pairwise surface gap > gap, uniform random
    rotations, v0 ~ N(0, vsigma^2) per axis (the 2000-body workload)."""
    rng = PCG32(seed)
    unit_v, f, e = icosphere_mesh(subdiv, 1.0)
    cell = 2 * rmax + gap
    grid = {}
    centers, radii = [], []
    tries = 0
    while len(centers) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("sphere packing failed; enlarge the box")
        r = rng.uniform(rmin, rmax)
        c = np.array([rng.uniform(-half + r, half - r), rng.uniform(ground_y + r + gap, ytop),
                      rng.uniform(-half + r, half - r)])
        key = tuple(np.floor(c / cell).astype(int))
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for j in grid.get((key[0] + dx, key[1] + dy, key[2] + dz), ()):
                        if np.linalg.norm(c - centers[j]) <= r + radii[j] + gap:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if not ok:
            continue
        grid.setdefault(key, []).append(len(centers))
        centers.append(c)
        radii.append(r)
    bodies = []
    for c, r in zip(centers, radii):
        R = rng.rotation()
        vel = (rng.normal(0, vsigma), rng.normal(0, vsigma), rng.normal(0, vsigma))
        bodies.append(dict(verts=unit_v * r, tris=f, edges=e, R=R, p=c, v=vel))
    return build_scene(bodies)


def torus_chain(n=64, R=0.1, r=0.025, seed=4, pin_top=True):
    """Chain of interlocked tori, alternate links rotated 90 degrees about the
    chain axis, top link static (the differentiable workload)."""
    rng = PCG32(seed)
    v, f, e = torus_mesh(R, r)
    d = 2 * R - 4 * r
    bodies = []
    for k in range(n):
        # torus local axis is z (ring in the x-y plane); alternate links rotate
        # 90 degrees about y so their ring lies in the z-y plane
        Rm = np.eye(3) if k % 2 == 0 else euler_rotation(0.0, math.pi / 2, 0.0)
        p = (1e-3 * rng.uniform(-1, 1), -d * k, 1e-3 * rng.uniform(-1, 1))
        m = -1000.0 if (pin_top and k == 0) else 1000.0
        bodies.append(dict(verts=v, tris=f, edges=e, R=Rm, p=p, mass=m, Ic=1000.0 / 12))
    return build_scene(bodies)


def card_lattice(n=10000, seed=5, size=(0.1, 0.002, 0.15), tilt_deg=12.0, pitch=0.12, ground_y=-0.5,
                 gap=0.05):
    """Thin boxes (cards) standing on their short edge in a lattice, alternate rows
    tilted by +/- tilt_deg (the leaning card-lattice workload)."""
    rng = PCG32(seed)
    v, f, e = cube_mesh((size[0], size[2], size[1]))   # x width, y height, z thickness
    side = int(math.ceil(math.sqrt(n)))
    bodies = []
    h = size[2]
    for k in range(n):
        i, j = divmod(k, side)
        sgn = 1.0 if (i % 2 == 0) else -1.0
        ang = math.radians(tilt_deg) * sgn + math.radians(1.0) * rng.uniform(-1, 1)
        Rm = euler_rotation(0.0, 0.0, ang)
        p = (j * pitch - side * pitch / 2, ground_y + 0.5 * h * math.cos(ang) + 0.5 * size[0] * abs(math.sin(ang)) + gap,
             i * (pitch + 0.1) - side * (pitch + 0.1) / 2)
        bodies.append(dict(verts=v, tris=f, edges=e, R=Rm, p=p))
    return build_scene(bodies)


def batched_scenes(n_scenes=256, seed=3, nmin=8, nmax=200, spacing=20.0, pitch=2.0, ground_y=-0.5):
    """Independent scenes of U(nmin, nmax) bodies, each a randomly rotated unit
    cube or a subdivision-2 icosphere (162 vertices, radius 0.5) with
    probability 1/2, stacked on a grid of `pitch` (more than a rotated cube's
    diagonal, sqrt(3)) above the ground, the lowest layer clear of it by
    0.034; the scenes sit `spacing` apart (batched stepping keeps them apart
    whatever their positions)."""
    rng = PCG32(seed)
    cv, cf, ce = cube_mesh()
    sv, sf, se = icosphere_mesh(2, 0.5)
    scenes = []
    side = int(math.ceil(math.sqrt(n_scenes)))
    y0 = ground_y + 0.5 * math.sqrt(3.0) + 0.034
    for s in range(n_scenes):
        nb = int(nmin + (nmax - nmin + 1) * rng.uniform())
        ox, oz = (s % side) * spacing, (s // side) * spacing
        bodies = []
        cols = int(math.ceil(nb ** (1 / 3)))
        for k in range(nb):
            a, b_, c = k % cols, (k // cols) % cols, k // (cols * cols)
            use_sphere = rng.uniform() < 0.5
            vv, ff, ee = (sv, sf, se) if use_sphere else (cv, cf, ce)
            p = (ox + pitch * a + rng.uniform(-0.05, 0.05), y0 + pitch * c, oz + pitch * b_ + rng.uniform(-0.05, 0.05))
            bodies.append(dict(verts=vv, tris=ff, edges=ee, R=rng.rotation(), p=p))
        scenes.append(bodies)
    flat = [b for sc in scenes for b in sc]
    return build_scene(flat), [len(sc) for sc in scenes]


def load_reference_scene_json(path_or_list, obj_dir=None, meshes=None):
    """Scenes in the reference's JSON format (test_cases/spinning_cube.cpp:56-127);
    see adb_amd.scene_io.load_scene_json."""
    from .scene_io import load_scene_json
    return load_scene_json(path_or_list, obj_dir=obj_dir, meshes=meshes)
