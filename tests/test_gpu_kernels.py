"""GPU parity of the kernel-level API (abdtk functions) against the oracle.

Bit-exact: distance types and squared distances (the contact predicate) and
CCD times of impact.  Tolerance (stated per test): barrier Hessians and
gradients, whose log() and eigen-decomposition differ from the host's by
rounding.  All calls go through the C ABI of adb_amd/libadb.so.
"""
import numpy as np
import pytest

import oracle_lib as O
from adb_amd import abdtk
from adb_amd.capi import kernel_config
from adb_amd.scenes import PCG32

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _require_gpu():
    assert torch.cuda.is_available(), "GPU tests need an MI355X"


def rand_configs(n, seed, parallel_frac=0.2):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, size=(n, 4, 3))
    # axis-aligned / exactly parallel edge pairs (the cube-stack case)
    k = int(n * parallel_frac)
    X[:k, 1] = X[:k, 0] + np.array([1.0, 0, 0])
    X[:k, 3] = X[:k, 2] + np.array([0.7, 0, 0])
    # snap some coordinates to a coarse grid to produce exact ties
    X[k:2 * k] = np.round(X[k:2 * k] * 4) / 4
    return X.reshape(n, 12)


def test_distance_bit_exact():
    _require_gpu()
    X = rand_configs(20000, 1)
    xt = torch.tensor(X, device="cuda")
    for kind, fn in (("pt", O.pt_distance), ("ee", O.ee_distance)):
        ty, d = abdtk.distance_batched(xt, kind)
        ty, d = ty.cpu().numpy(), d.cpu().numpy()
        ref = [fn(x) for x in X]
        rt = np.array([r[0] for r in ref])
        rd = np.array([r[1] for r in ref])
        assert np.array_equal(ty, rt), (kind, np.nonzero(ty != rt)[0][:10])
        both_nan = np.isnan(d) & np.isnan(rd)
        assert np.array_equal(d[~both_nan], rd[~both_nan]), kind


def test_collision_time_bit_exact():
    _require_gpu()
    rng = np.random.default_rng(2)
    n = 8000
    P0 = rng.uniform(0, 1, size=(n, 4, 3))
    P1 = P0 + rng.uniform(-1, 1, size=(n, 4, 3))
    # deliberate PT collisions (Sample-Test1/test.cpp:9-31)
    toi = rng.uniform(0.01, 1, size=n // 4)
    P1[: n // 4] = P0[: n // 4]
    P1[: n // 4, 0] = P0[: n // 4, 2] / toi[:, None] - P0[: n // 4, 0] * (1 - toi[:, None]) / toi[:, None]
    X = np.concatenate([P0.reshape(n, 12), P1.reshape(n, 12)], axis=1)
    xt = torch.tensor(X, device="cuda")
    for kind, fn in (("pt", O.pt_collision_time), ("ee", O.ee_collision_time)):
        got = abdtk.collision_time_batched(xt, kind).cpu().numpy()
        ref = np.array([fn(x) for x in X])
        assert np.array_equal(got, ref), (kind, np.nonzero(got != ref)[0][:10])
        assert (got < 1).sum() > n // 8


def near_contacts(n, seed, kind):
    """Configurations with squared distance below d_hat = 1e-2."""
    rng = PCG32(seed)
    out, types, dists = [], [], []
    while len(out) < n:
        P = np.array([[rng.uniform(-0.5, 0.5) for _ in range(3)] for _ in range(4)])
        if kind == "pt":
            P[0] = P[1:].mean(0) + 0.08 * np.array([rng.uniform(-1, 1) for _ in range(3)])
            t, d = O.pt_distance(P.reshape(-1))
        else:
            m = 0.5 * (P[2] + P[3])
            P[0], P[1] = m + (P[0] - P[1]) * 0.5, m - (P[0] - P[1]) * 0.5
            P[0] += 0.05 * np.array([rng.uniform(-1, 1) for _ in range(3)])
            t, d = O.ee_distance(P.reshape(-1))
        if 1e-8 < d < 1e-2:
            out.append(P.reshape(-1)); types.append(t); dists.append(d)
    return np.array(out), np.array(types, dtype=np.int32), np.array(dists)


@pytest.mark.parametrize("kind", ["pt", "ee"])
@pytest.mark.parametrize("psd", [0, 1])
def test_ipc_hess_batched_vs_oracle(kind, psd):
    _require_gpu()
    cfg = kernel_config(d_hat=1e-2, kappa=1e-3, eps_x=1e-3, psd=psd)
    X, T, D = near_contacts(400, 7 if kind == "pt" else 8, kind)
    old = abdtk.config
    abdtk.config = cfg
    try:
        xt, tt, dt = (torch.tensor(X, device="cuda"), torch.tensor(T, device="cuda"), torch.tensor(D, device="cuda"))
        if kind == "pt":
            H, g = abdtk.ipc_hess_pt_12x12_batched(xt, tt, dt)
        else:
            H, g, m = abdtk.ipc_hess_ee_12x12_batched(xt, tt, dt)
            m = m.cpu().numpy()
        H, g = H.cpu().numpy(), g.cpu().numpy()
    finally:
        abdtk.config = old
    for i in range(len(X)):
        if kind == "pt":
            Hr, gr = O.ipc_hess_pt(cfg, X[i], int(T[i]), D[i])
        else:
            Hr, gr, mr = O.ipc_hess_ee(cfg, X[i], int(T[i]), D[i])
            assert m[i] == pytest.approx(mr, rel=1e-13, abs=1e-15)
        scale = max(np.abs(Hr).max(), 1e-300)
        # tolerance: 1e-9 of the largest entry (log() and Jacobi vs host rounding)
        np.testing.assert_allclose(H[i], Hr, rtol=0, atol=1e-9 * scale)
        np.testing.assert_allclose(g[i], gr, rtol=1e-10, atol=1e-12 * max(np.abs(gr).max(), 1e-300))


def test_abdtk_reference_example():
    """The input of python/test_bindings.py:5-20 through the abdtk mirror."""
    _require_gpu()
    arr = np.array([[0.4360, 0.0259, 0.5497], [0.4353, 0.4204, 0.3303], [0.2046, 0.6193, 0.2997],
                    [0.2668, 0.6211, 0.5291]])
    t, d = O.pt_distance(arr.reshape(-1))
    cfg = abdtk.config
    H, g = abdtk.ipc_hess_pt_12x12([arr[i] for i in range(4)], list(range(4)), abdtk.PointTriangleDistanceType(t), d)
    Hr, gr = O.ipc_hess_pt(cfg, arr.reshape(-1), t, d)
    np.testing.assert_allclose(H, Hr, rtol=0, atol=1e-9 * max(np.abs(Hr).max(), 1e-300))
    np.testing.assert_allclose(g, gr, rtol=1e-10, atol=1e-300)
    x = np.concatenate([arr.reshape(-1), arr.reshape(-1) + 0.1])
    assert abdtk.pt_collision_time(*x.reshape(8, 3)) == O.pt_collision_time(x)
    assert abdtk.ee_collision_time(*x.reshape(8, 3)) == O.ee_collision_time(x)
