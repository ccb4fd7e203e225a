"""The no-allocation collision-time entry point (adb_collision_time_noalloc)
against the entry points it varies: the single-entry host-pointer
adb_pt/ee_collision_time and the batched device-pointer forms, bit for bit,
on two calls in a row (the second smaller), and its size limit."""
import ctypes as C

import numpy as np
import pytest

from adb_amd import capi

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NOALLOC_MAX = 4096


def ccd_inputs(n, seed):
    """n entries of 4 points at t0 and t1: random points moving towards each
    other, so that a share of them collide within the step."""
    rng = np.random.default_rng(seed)
    x0 = rng.uniform(-1, 1, size=(n, 4, 3))
    x0[:, 1:] = x0[:, :1] + rng.uniform(-0.3, 0.3, size=(n, 3, 3))
    v = rng.normal(0, 0.4, size=(n, 4, 3))
    return np.ascontiguousarray(np.concatenate([x0, x0 + v], axis=1).reshape(n, 24))


def noalloc(lib, is_ee, x):
    n = len(x)
    toi = np.full(n, np.nan)
    rc = lib.adb_collision_time_noalloc(is_ee, n, capi.dptr(x), capi.dptr(toi))
    return rc, toi


@pytest.mark.parametrize("is_ee", [0, 1])
def test_noalloc_matches_existing_entry_points(is_ee):
    assert torch.cuda.is_available(), "GPU tests need an MI355X"
    lib = capi.load_library()
    single = lib.adb_ee_collision_time if is_ee else lib.adb_pt_collision_time
    batched = lib.adb_ee_collision_time_batched if is_ee else lib.adb_pt_collision_time_batched
    for n, seed in ((3000, 1 + is_ee), (257, 7 + is_ee)):      # second call smaller
        x = ccd_inputs(n, seed)
        rc, toi = noalloc(lib, is_ee, x)
        assert rc == capi.ADB_OK
        # batched device-pointer form on the same inputs
        xd = torch.from_numpy(x).to("cuda:0")
        td = torch.empty(n, dtype=torch.float64, device="cuda:0")
        assert batched(n, C.c_void_p(xd.data_ptr()), C.c_void_p(td.data_ptr()), None) == capi.ADB_OK
        torch.cuda.synchronize()
        ref = td.cpu().numpy()
        assert np.array_equal(toi, ref)
        # single-entry host form on a sample
        for i in range(0, n, max(1, n // 64)):
            assert single(capi.dptr(np.ascontiguousarray(x[i]))) == toi[i], (n, i)
        assert np.sum(toi < 1.0) > n // 20 and np.sum(toi >= 1.0) > 0


def test_noalloc_rejects_beyond_limit():
    lib = capi.load_library()
    x = np.zeros((NOALLOC_MAX + 1, 24))
    toi = np.zeros(NOALLOC_MAX + 1)
    assert lib.adb_collision_time_noalloc(0, NOALLOC_MAX + 1, capi.dptr(x), capi.dptr(toi)) == capi.ADB_E_CAPACITY
    assert lib.adb_collision_time_noalloc(2, 1, capi.dptr(x), capi.dptr(toi)) == capi.ADB_E_ARG
    assert lib.adb_collision_time_noalloc(0, 0, None, None) == capi.ADB_OK
