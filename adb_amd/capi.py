"""ctypes view of the C ABI declared in include/adb_capi.h.

The structs mirror the header field for field.  `load_library()` loads the
in-tree HIP library (adb_amd/libadb.so) and raises if it is missing: there is
no CPU fallback anywhere in the product path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libadb.so")

ADB_OK = 0
ADB_E_ARG = -1
ADB_E_HIP = -2
ADB_E_NAN = -3
ADB_E_TOI_STALL = -4
ADB_E_CAPACITY = -5
ADB_E_NOT_CONVERGED = -6
ADB_E_NO_DEVICE = -7
ADB_E_SOLVER = -8

STATUS_NAMES = {
    ADB_OK: "ADB_OK", ADB_E_ARG: "ADB_E_ARG", ADB_E_HIP: "ADB_E_HIP", ADB_E_NAN: "ADB_E_NAN",
    ADB_E_TOI_STALL: "ADB_E_TOI_STALL", ADB_E_CAPACITY: "ADB_E_CAPACITY",
    ADB_E_NOT_CONVERGED: "ADB_E_NOT_CONVERGED", ADB_E_NO_DEVICE: "ADB_E_NO_DEVICE",
    ADB_E_SOLVER: "ADB_E_SOLVER",
}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class AdbScene(C.Structure):
    _fields_ = [
        ("n_bodies", C.c_int),
        ("vert_offset", _ip), ("face_offset", _ip), ("edge_offset", _ip),
        ("rest_verts", _dp), ("faces", _ip), ("edges", _ip),
        ("mass", _dp), ("Ic", _dp), ("q0", _dp), ("dqdt", _dp),
    ]


class AdbParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("d_hat", C.c_double), ("kappa", C.c_double),
        ("kappa_ortho", C.c_double), ("mu", C.c_double), ("evh", C.c_double),
        ("eps_x", C.c_double), ("tol", C.c_double), ("c1", C.c_double),
        ("backoff", C.c_double), ("max_uk", C.c_double),
        ("gravity", C.c_double * 3), ("ground_y", C.c_double),
        ("max_iter", C.c_int), ("max_ls_iter", C.c_int),
        ("enable_pt", C.c_int), ("enable_ee", C.c_int), ("enable_ground", C.c_int),
        ("pt_fric", C.c_int), ("ee_fric", C.c_int), ("vg_fric", C.c_int),
        ("psd", C.c_int), ("line_search", C.c_int), ("upper_bound", C.c_int),
        ("broad_phase_fp32", C.c_int),
        ("pcg_rtol", C.c_double), ("pcg_max_iter", C.c_int),
        ("iaabb", C.c_int), ("sh_bits", C.c_int), ("sh_min_xyz", C.c_double), ("sh_dx", C.c_double),
    ]


class AdbStepStats(C.Structure):
    _fields_ = [
        ("newton_iters", C.c_int), ("ls_trials", C.c_int), ("pcg_iters", C.c_int),
        ("n_pt", C.c_int), ("n_ee", C.c_int), ("n_vg", C.c_int),
        ("contacts_processed", C.c_longlong),
        ("toi_min", C.c_double),
        ("t_ipc", C.c_double), ("t_solver", C.c_double), ("t_ccd", C.c_double),
        ("t_line_search", C.c_double), ("t_total", C.c_double),
        ("pcg_not_converged", C.c_int), ("pcg_breakdown", C.c_int),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class AdbTuning(C.Structure):
    _fields_ = [("narrow_per_pair", C.c_double), ("ee_slices", C.c_int), ("pcg_batch", C.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class AdbKernelConfig(C.Structure):
    _fields_ = [("d_hat", C.c_double), ("kappa", C.c_double), ("eps_x", C.c_double), ("psd", C.c_int)]


def default_params(**overrides):
    """The reference's defaults: config.json, model/settings.h, model/barrier.cpp:7-8,
    model/othogonal_energy.cpp:8, model/geometry.cpp:60.  Same values as
    adb_params_default() in the library."""
    p = AdbParams()
    p.dt = 1e-2
    p.d_hat = 1e-4
    p.kappa = 1e-4
    p.kappa_ortho = 1e9
    p.mu = 0.0
    p.evh = 1e-3
    p.eps_x = 1e-3
    p.tol = 1e-2
    p.c1 = 1e-4
    p.backoff = 0.8
    p.max_uk = 0.0
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, -9.8, 0.0
    p.ground_y = -0.5
    p.max_iter = 1000
    p.max_ls_iter = 200
    p.enable_pt = p.enable_ee = p.enable_ground = 1
    p.pt_fric = p.ee_fric = p.vg_fric = 1
    p.psd = 1
    p.line_search = 1
    p.upper_bound = 1
    p.broad_phase_fp32 = 0
    p.pcg_rtol = 1e-12
    p.pcg_max_iter = 2000
    p.iaabb = 3
    p.sh_bits = 6
    p.sh_min_xyz = -10.0
    p.sh_dx = 0.8
    for k, v in overrides.items():
        if k == "gravity":
            for i in range(3):
                p.gravity[i] = v[i]
        else:
            setattr(p, k, v)
    return p


def kernel_config(d_hat=1e-4, kappa=1e-3, eps_x=1e-3, psd=1):
    """abdtk's frozen constants: barrier::d_hat = 1e-4 and barrier::kappa = 1e-3
    (model/barrier.cpp:7-8, never reassigned in the abdtk build), globals.psd = true
    and globals.eps_x = 1e-3 (model/settings.h:16,39)."""
    c = AdbKernelConfig()
    c.d_hat, c.kappa, c.eps_x, c.psd = d_hat, kappa, eps_x, psd
    return c


class SceneArrays:
    """Owns the numpy arrays behind an AdbScene (keeps them alive)."""

    def __init__(self, scene):
        self.np = {
            "vert_offset": np.ascontiguousarray(scene["vert_offset"], dtype=np.int32),
            "face_offset": np.ascontiguousarray(scene["face_offset"], dtype=np.int32),
            "edge_offset": np.ascontiguousarray(scene["edge_offset"], dtype=np.int32),
            "rest_verts": np.ascontiguousarray(scene["rest_verts"], dtype=np.float64).reshape(-1),
            "faces": np.ascontiguousarray(scene["faces"], dtype=np.int32).reshape(-1),
            "edges": np.ascontiguousarray(scene["edges"], dtype=np.int32).reshape(-1),
            "mass": np.ascontiguousarray(scene["mass"], dtype=np.float64),
            "Ic": np.ascontiguousarray(scene["Ic"], dtype=np.float64),
            "q0": np.ascontiguousarray(scene["q0"], dtype=np.float64).reshape(-1),
            "dqdt": np.ascontiguousarray(scene["dqdt"], dtype=np.float64).reshape(-1),
        }
        s = AdbScene()
        s.n_bodies = len(self.np["mass"])
        for k, a in self.np.items():
            ptr_t = _ip if a.dtype == np.int32 else _dp
            setattr(s, k, a.ctypes.data_as(ptr_t))
        self.struct = s


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


_LIB = None


def load_library():
    """Load the HIP product library.  Raises when it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "adb_amd/libadb.so is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(hipcc --offload-arch=gfx950). There is no CPU fallback.")
    lib = C.CDLL(LIB_PATH)
    _declare(lib)
    _LIB = lib
    return lib


def _declare(lib):
    vp = C.c_void_p
    sig = {
        "adb_params_default": (None, [C.POINTER(AdbParams)]),
        "adb_kernel_config_default": (None, [C.POINTER(AdbKernelConfig)]),
        "adb_world_create": (C.c_int, [C.POINTER(AdbScene), C.POINTER(AdbParams), C.POINTER(vp)]),
        "adb_world_destroy": (None, [vp]),
        "adb_world_create_batched": (C.c_int, [C.POINTER(AdbScene), _ip, C.c_int, C.POINTER(AdbParams), C.POINTER(vp)]),
        "adb_batch_stats": (C.c_int, [vp, _ip, _ip, _ip, _ip]),
        "adb_step": (C.c_int, [vp, C.POINTER(AdbStepStats)]),
        "adb_get_state": (C.c_int, [vp, _dp, _dp]),
        "adb_set_state": (C.c_int, [vp, _dp, _dp]),
        "adb_last_error": (C.c_char_p, [vp]),
        "adb_collision_set": (C.c_int, [vp, C.c_int, _dp, _ip, C.c_int, _ip, _ip, C.c_int, _ip, _ip, C.c_int, _ip]),
        "adb_step_size_upper_bound": (C.c_int, [vp, _dp, _dp]),
        "adb_assemble": (C.c_int, [vp, _dp, _dp]),
        "adb_newton_direction": (C.c_int, [vp, _dp]),
        "adb_contact_terms": (C.c_int, [vp, _dp, _dp, _dp]),
        "adb_adjoint_solve": (C.c_int, [vp, _dp, _dp, _ip]),
        "adb_line_search": (C.c_int, [vp, _dp, _dp, _dp, _dp, _dp, _ip]),
        "adb_ipc_hess_pt_12x12_batched": (C.c_int, [C.POINTER(AdbKernelConfig), C.c_int, vp, vp, vp, vp, vp, vp]),
        "adb_ipc_hess_ee_12x12_batched": (C.c_int, [C.POINTER(AdbKernelConfig), C.c_int, vp, vp, vp, vp, vp, vp, vp]),
        "adb_pt_collision_time_batched": (C.c_int, [C.c_int, vp, vp, vp]),
        "adb_ee_collision_time_batched": (C.c_int, [C.c_int, vp, vp, vp]),
        "adb_ipc_hess_pt_12x12": (C.c_int, [C.POINTER(AdbKernelConfig), _dp, C.c_int, C.c_double, _dp, _dp]),
        "adb_ipc_hess_ee_12x12": (C.c_int, [C.POINTER(AdbKernelConfig), _dp, C.c_int, C.c_double, _dp, _dp, _dp]),
        "adb_pt_collision_time": (C.c_double, [_dp]),
        "adb_ee_collision_time": (C.c_double, [_dp]),
        "adb_pt_distance_batched": (C.c_int, [C.c_int, vp, vp, vp, vp]),
        "adb_ee_distance_batched": (C.c_int, [C.c_int, vp, vp, vp, vp]),
        "adb_device_count": (C.c_int, []),
        "adb_collision_time_noalloc": (C.c_int, [C.c_int, C.c_int, _dp, _dp]),
        "adb_profile_enable": (C.c_int, [vp, C.c_int]),
        "adb_profile_read": (C.c_int, [vp, C.c_int, _dp, C.POINTER(C.c_longlong), _dp]),
        "adb_profile_flops": (C.c_int, [vp, C.c_int, _dp]),
        "adb_get_tuning": (C.c_int, [vp, C.POINTER(AdbTuning)]),
        "adb_set_tuning": (C.c_int, [vp, C.POINTER(AdbTuning)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


# Every symbol include/adb_capi.h declares (checked by tests/test_capi_symbols.py).
EXPORTED_SYMBOLS = [
    "adb_params_default", "adb_world_create", "adb_world_create_batched", "adb_batch_stats", "adb_world_destroy",
    "adb_step", "adb_get_state",
    "adb_set_state", "adb_last_error", "adb_collision_set", "adb_step_size_upper_bound",
    "adb_assemble", "adb_newton_direction", "adb_contact_terms", "adb_line_search", "adb_adjoint_solve",
    "adb_kernel_config_default",
    "adb_ipc_hess_pt_12x12_batched", "adb_ipc_hess_ee_12x12_batched",
    "adb_pt_collision_time_batched", "adb_ee_collision_time_batched",
    "adb_ipc_hess_pt_12x12", "adb_ipc_hess_ee_12x12", "adb_pt_collision_time",
    "adb_ee_collision_time", "adb_pt_distance_batched", "adb_ee_distance_batched",
    "adb_device_count", "adb_profile_enable", "adb_profile_read", "adb_profile_flops", "adb_get_tuning", "adb_set_tuning", "adb_collision_time_noalloc",
]

# kernel ids of adb_profile_read (enum adb_prof_kernel)
PROF_KERNELS = ["transform", "body_box", "body_pairs", "narrow_set", "vg_scan", "sort", "narrow_ccd", "vg_toi",
                "contact", "vg_contact", "inertia", "asm_reduce", "block_chol", "spmv", "pcg_update", "pcg_dir",
                "energy", "contact_classify", "contact_group", "hash_build", "hash_query", "contact_pk"]
