"""`abdtk` - the reference's Python operator API for the contact kernels
(python/bindings.cpp:42-83), served by the gfx950 kernels of libadb.so.

Same names, enum values and argument meaning as the reference module:

    ipc_hess_pt_12x12(pt, ij, pt_type, dist) -> (H 12x12, g 12)
    ipc_hess_ee_12x12(ee, ij, ee_type, dist) -> (H 12x12, g 12, mollifier)
    pt_collision_time(p0_t0, t0_t0, t1_t0, t2_t0, p0_t1, t0_t1, t1_t1, t2_t1) -> toi
    ee_collision_time(ei0_t0, ei1_t0, ej0_t0, ej1_t0, ei0_t1, ei1_t1, ej0_t1, ej1_t1) -> toi

`dist` is the SQUARED distance (model/barrier.cpp:7-8).  Constants are the
abdtk build's frozen globals (barrier::d_hat = 1e-4, barrier::kappa = 1e-3,
psd = true, eps_x = 1e-3); `config` can override them.  Batched variants take
torch tensors already on the GPU.  There is no CPU path: if the HIP library
is missing or no gfx950 device is present, calls raise.
"""
import ctypes as C
import enum

import numpy as np

from .capi import ADB_OK, dptr, kernel_config, load_library


class PointTriangleDistanceType(enum.IntEnum):
    P_T0 = 0
    P_T1 = 1
    P_T2 = 2
    P_E0 = 3
    P_E1 = 4
    P_E2 = 5
    P_T = 6
    AUTO = 7


class EdgeEdgeDistanceType(enum.IntEnum):
    EA0_EB0 = 0
    EA0_EB1 = 1
    EA1_EB0 = 2
    EA1_EB1 = 3
    EA_EB0 = 4
    EA_EB1 = 5
    EA0_EB = 6
    EA1_EB = 7
    EA_EB = 8
    AUTO = 9


# export_values() in the reference binding
globals().update({m.name: m for m in PointTriangleDistanceType if m.name != "AUTO"})
globals().update({m.name: m for m in EdgeEdgeDistanceType if m.name != "AUTO"})

config = kernel_config()


def _pts(pts):
    a = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).reshape(-1))
    if a.size != 12:
        raise ValueError("expected 4 points (12 coordinates)")
    return a


def ipc_hess_pt_12x12(pt, ij, pt_type, dist):
    lib = load_library()
    x = _pts(pt)
    H = np.zeros(144)
    g = np.zeros(12)
    rc = lib.adb_ipc_hess_pt_12x12(C.byref(config), dptr(x), int(pt_type), float(dist), dptr(H), dptr(g))
    if rc != ADB_OK:
        raise RuntimeError("adb_ipc_hess_pt_12x12 failed: %d" % rc)
    return H.reshape(12, 12).T.copy(), g


def ipc_hess_ee_12x12(ee, ij, ee_type, dist):
    lib = load_library()
    x = _pts(ee)
    H = np.zeros(144)
    g = np.zeros(12)
    m = C.c_double()
    rc = lib.adb_ipc_hess_ee_12x12(C.byref(config), dptr(x), int(ee_type), float(dist), dptr(H), dptr(g), C.byref(m))
    if rc != ADB_OK:
        raise RuntimeError("adb_ipc_hess_ee_12x12 failed: %d" % rc)
    return H.reshape(12, 12).T.copy(), g, m.value


def pt_collision_time(p0_t0, t0_t0, t1_t0, t2_t0, p0_t1, t0_t1, t1_t1, t2_t1):
    x = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.float64).reshape(3) for v in
                                             (p0_t0, t0_t0, t1_t0, t2_t0, p0_t1, t0_t1, t1_t1, t2_t1)]))
    t = load_library().adb_pt_collision_time(dptr(x))
    if t < 0:
        raise RuntimeError("adb_pt_collision_time failed")
    return t


def ee_collision_time(ei0_t0, ei1_t0, ej0_t0, ej1_t0, ei0_t1, ei1_t1, ej0_t1, ej1_t1):
    x = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.float64).reshape(3) for v in
                                             (ei0_t0, ei1_t0, ej0_t0, ej1_t0, ei0_t1, ei1_t1, ej0_t1, ej1_t1)]))
    t = load_library().adb_ee_collision_time(dptr(x))
    if t < 0:
        raise RuntimeError("adb_ee_collision_time failed")
    return t


# ---------------------------------------------------------------- batched (torch, on device)
def _dev(t, dtype):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda):
        raise ValueError("batched abdtk calls take GPU tensors")
    return t.to(dtype).contiguous()


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def ipc_hess_pt_12x12_batched(pt, pt_type, dist):
    """pt: (N, 12) float64, pt_type: (N,) int32, dist: (N,) -> H (N, 12, 12), g (N, 12)."""
    import torch
    lib = load_library()
    pt = _dev(pt, torch.float64)
    ty = _dev(pt_type, torch.int32)
    d = _dev(dist, torch.float64)
    n = pt.shape[0]
    H = torch.empty((n, 12, 12), dtype=torch.float64, device=pt.device)
    g = torch.empty((n, 12), dtype=torch.float64, device=pt.device)
    rc = lib.adb_ipc_hess_pt_12x12_batched(C.byref(config), n, C.c_void_p(pt.data_ptr()), C.c_void_p(ty.data_ptr()),
                                           C.c_void_p(d.data_ptr()), C.c_void_p(H.data_ptr()), C.c_void_p(g.data_ptr()),
                                           _stream())
    if rc != ADB_OK:
        raise RuntimeError("adb_ipc_hess_pt_12x12_batched failed: %d" % rc)
    return H.transpose(1, 2), g     # stored column-major


def ipc_hess_ee_12x12_batched(ee, ee_type, dist):
    import torch
    lib = load_library()
    ee = _dev(ee, torch.float64)
    ty = _dev(ee_type, torch.int32)
    d = _dev(dist, torch.float64)
    n = ee.shape[0]
    H = torch.empty((n, 12, 12), dtype=torch.float64, device=ee.device)
    g = torch.empty((n, 12), dtype=torch.float64, device=ee.device)
    m = torch.empty((n,), dtype=torch.float64, device=ee.device)
    rc = lib.adb_ipc_hess_ee_12x12_batched(C.byref(config), n, C.c_void_p(ee.data_ptr()), C.c_void_p(ty.data_ptr()),
                                           C.c_void_p(d.data_ptr()), C.c_void_p(H.data_ptr()), C.c_void_p(g.data_ptr()),
                                           C.c_void_p(m.data_ptr()), _stream())
    if rc != ADB_OK:
        raise RuntimeError("adb_ipc_hess_ee_12x12_batched failed: %d" % rc)
    return H.transpose(1, 2), g, m


def collision_time_batched(x, kind="pt"):
    """x: (N, 24) float64 (4 points at t0 then 4 at t1) -> toi (N,)."""
    import torch
    lib = load_library()
    x = _dev(x, torch.float64)
    n = x.shape[0]
    t = torch.empty((n,), dtype=torch.float64, device=x.device)
    f = lib.adb_pt_collision_time_batched if kind == "pt" else lib.adb_ee_collision_time_batched
    rc = f(n, C.c_void_p(x.data_ptr()), C.c_void_p(t.data_ptr()), _stream())
    if rc != ADB_OK:
        raise RuntimeError("collision_time_batched failed: %d" % rc)
    return t


def distance_batched(x, kind="pt"):
    """x: (N, 12) float64 -> (type int32 (N,), squared distance (N,))."""
    import torch
    lib = load_library()
    x = _dev(x, torch.float64)
    n = x.shape[0]
    ty = torch.empty((n,), dtype=torch.int32, device=x.device)
    d = torch.empty((n,), dtype=torch.float64, device=x.device)
    f = lib.adb_pt_distance_batched if kind == "pt" else lib.adb_ee_distance_batched
    rc = f(n, C.c_void_p(x.data_ptr()), C.c_void_p(ty.data_ptr()), C.c_void_p(d.data_ptr()), _stream())
    if rc != ADB_OK:
        raise RuntimeError("distance_batched failed: %d" % rc)
    return ty, d
