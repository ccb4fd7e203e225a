"""The pybind11 `abdtk` module (adb_amd/pyext, built from
adb_amd/csrc/abdtk_pybind.cpp over libadb.so) against the reference module's
interface (python/bindings.cpp:42-83) and, on the GPU, against the ctypes
binding and the oracle.

CPU: the module imports, exports the reference's names, enum values and
argument names.  GPU: the reference's own example (python/test_bindings.py:5-18,
point-plane contact) and random contacts give the same H, g, mollifier and
collision times as the ctypes binding bit for bit, and the oracle within 1e-9.
"""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from adb_amd import abdtk as ct
from adb_amd.capi import kernel_config

PYEXT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "adb_amd", "pyext")


@pytest.fixture(scope="module")
def abdtk():
    sys.path.insert(0, PYEXT)
    try:
        import abdtk as m
    finally:
        sys.path.remove(PYEXT)
    assert m.__file__.startswith(PYEXT)
    return m


def test_module_interface(abdtk):
    pt_names = ["P_T0", "P_T1", "P_T2", "P_E0", "P_E1", "P_E2", "P_T", "AUTO"]
    ee_names = ["EA0_EB0", "EA0_EB1", "EA1_EB0", "EA1_EB1", "EA_EB0", "EA_EB1", "EA0_EB", "EA1_EB", "EA_EB", "AUTO"]
    for i, n in enumerate(pt_names):
        assert int(getattr(abdtk.PointTriangleDistanceType, n)) == i
    for i, n in enumerate(ee_names):
        assert int(getattr(abdtk.EdgeEdgeDistanceType, n)) == i
    assert int(abdtk.P_T) == 6 and int(abdtk.EA_EB) == 8          # export_values()
    doc = abdtk.ipc_hess_pt_12x12.__doc__
    assert "pt" in doc and "ij" in doc and "pt_type" in doc and "dist" in doc
    assert "ee_type" in abdtk.ipc_hess_ee_12x12.__doc__
    assert "p0_t0" in abdtk.pt_collision_time.__doc__ and "ej1_t1" in abdtk.ee_collision_time.__doc__


@pytest.mark.gpu
def test_reference_example_and_random_contacts(abdtk):
    torch = pytest.importorskip("torch")
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    arr = np.array([[0.4360, 0.0259, 0.5497], [0.4353, 0.4204, 0.3303],
                    [0.2046, 0.6193, 0.2997], [0.2668, 0.6211, 0.5291]])
    cfg = kernel_config()
    # as the example: type P_T with the squared point-plane distance (there
    # ipctk.point_plane_distance); the example's distance is beyond d_hat, so
    # the configuration is also scaled into the barrier band
    for scale in (1.0, 0.01):
        P = arr * scale
        n = np.cross(P[2] - P[1], P[3] - P[1])
        d = float(np.dot(P[0] - P[1], n) ** 2 / np.dot(n, n))
        H, g = abdtk.ipc_hess_pt_12x12([P[i] for i in range(4)], list(range(4)), abdtk.P_T, d)
        Hc, gc = ct.ipc_hess_pt_12x12(P, None, 6, d)
        assert np.array_equal(H, Hc) and np.array_equal(g, gc)
        Ho, go = O.ipc_hess_pt(cfg, P.reshape(-1), 6, d)
        if d < cfg.d_hat:
            assert np.abs(Ho).max() > 0
        np.testing.assert_allclose(H, Ho, rtol=0, atol=1e-9 * max(1e-300, np.abs(Ho).max()))
        np.testing.assert_allclose(g, go, rtol=0, atol=1e-9 * max(1e-300, np.abs(go).max()))
    rng = np.random.default_rng(3)
    for _ in range(50):
        P = rng.uniform(-1, 1, size=(4, 3)) * 0.01
        te, de = O.ee_distance(P.reshape(-1))
        H, g, m = abdtk.ipc_hess_ee_12x12(P, None, abdtk.EdgeEdgeDistanceType(te), de)
        Hc, gc, mc = ct.ipc_hess_ee_12x12(P, None, te, de)
        assert np.array_equal(H, Hc) and np.array_equal(g, gc) and m == mc
        X = rng.uniform(-1, 1, size=(8, 3))
        X[4:] = X[:4] + rng.normal(0, 0.5, size=(4, 3))
        assert abdtk.pt_collision_time(*X) == ct.pt_collision_time(*X) == O.pt_collision_time(X.reshape(-1))
        assert abdtk.ee_collision_time(*X) == ct.ee_collision_time(*X) == O.ee_collision_time(X.reshape(-1))
