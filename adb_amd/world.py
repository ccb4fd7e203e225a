"""Python handle on an adb_world (include/adb_capi.h) - the MI355X Newton
contact step.  Mirrors the reference's ABD driver (model/abd.h:9-80): one
`step()` is ABD::implicit_euler (model/abd.cpp:30-140)."""
import ctypes as C

import numpy as np

from .capi import (ADB_OK, STATUS_NAMES, AdbStepStats, SceneArrays, default_params, dptr, iptr,
                   load_library)


class AdbError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (STATUS_NAMES.get(code, "?"), code, msg))
        self.code = code


class World:
    def __init__(self, scene, params=None, scene_offsets=None):
        """One world over `scene`; with scene_offsets (n_scenes + 1 body
        offsets) the bodies form independent scenes stepped together
        (adb_world_create_batched)."""
        self.lib = load_library()
        self.params = params if params is not None else default_params()
        self.arrays = SceneArrays(scene)
        self.B = len(scene["mass"])
        h = C.c_void_p()
        if scene_offsets is None:
            self.n_scenes = 1
            rc = self.lib.adb_world_create(C.byref(self.arrays.struct), C.byref(self.params), C.byref(h))
        else:
            self._soff = np.ascontiguousarray(scene_offsets, dtype=np.int32)
            self.n_scenes = len(self._soff) - 1
            rc = self.lib.adb_world_create_batched(C.byref(self.arrays.struct), iptr(self._soff), self.n_scenes,
                                                   C.byref(self.params), C.byref(h))
        if rc != ADB_OK:
            raise AdbError(rc, "adb_world_create failed")
        self.h = h

    def tuning(self):
        from .capi import AdbTuning
        t = AdbTuning()
        self._check(self.lib.adb_get_tuning(self.h, C.byref(t)), "adb_get_tuning")
        return t.as_dict()

    def set_tuning(self, narrow_per_pair=None, ee_slices=None, pcg_batch=None):
        from .capi import AdbTuning
        t = AdbTuning()
        self._check(self.lib.adb_get_tuning(self.h, C.byref(t)), "adb_get_tuning")
        if narrow_per_pair is not None:
            t.narrow_per_pair = narrow_per_pair
        if ee_slices is not None:
            t.ee_slices = ee_slices
        if pcg_batch is not None:
            t.pcg_batch = pcg_batch
        self._check(self.lib.adb_set_tuning(self.h, C.byref(t)), "adb_set_tuning")

    def batch_stats(self):
        """Per-scene (newton_iters, ls_trials, status, contacts) of the last step."""
        out = [np.zeros(self.n_scenes, dtype=np.int32) for _ in range(4)]
        self._check(self.lib.adb_batch_stats(self.h, *[iptr(a) for a in out]), "adb_batch_stats")
        return tuple(out)

    def close(self):
        if getattr(self, "h", None) is not None:
            self.lib.adb_world_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what, allow=()):
        if rc != ADB_OK and rc not in allow:
            raise AdbError(rc, "%s: %s" % (what, self.lib.adb_last_error(self.h).decode()))
        return rc

    def step(self, allow_not_converged=False):
        """One ABD::implicit_euler step.  Raises AdbError on any status other
        than ADB_OK (ADB_E_NOT_CONVERGED too, unless allow_not_converged)."""
        st = AdbStepStats()
        rc = self.lib.adb_step(self.h, C.byref(st))
        allow = (-6,) if allow_not_converged else ()
        self._check(rc, "adb_step", allow)
        return rc, st

    def state(self):
        q = np.zeros(12 * self.B)
        qd = np.zeros(12 * self.B)
        self._check(self.lib.adb_get_state(self.h, dptr(q), dptr(qd)), "adb_get_state")
        return q.reshape(-1, 12), qd.reshape(-1, 12)

    def set_state(self, q, qdot=None):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        qd = None if qdot is None else np.ascontiguousarray(qdot, dtype=np.float64).reshape(-1)
        self._check(self.lib.adb_set_state(self.h, dptr(q), dptr(qd) if qd is not None else None), "adb_set_state")

    def collision_set(self, vtn=1, dq=None, cap=1 << 20):
        while True:
            pt = np.zeros(4 * cap, dtype=np.int32)
            ee = np.zeros(4 * cap, dtype=np.int32)
            vg = np.zeros(2 * cap, dtype=np.int32)
            npt, nee, nvg = C.c_int(), C.c_int(), C.c_int()
            dqa = None if dq is None else np.ascontiguousarray(dq, dtype=np.float64).reshape(-1)
            rc = self.lib.adb_collision_set(self.h, vtn, dptr(dqa) if dqa is not None else None, iptr(pt), cap,
                                            C.byref(npt), iptr(ee), cap, C.byref(nee), iptr(vg), cap, C.byref(nvg))
            if rc == -5:
                cap = max(npt.value, nee.value, nvg.value) + 16
                continue
            self._check(rc, "adb_collision_set")
            return (pt[:4 * npt.value].reshape(-1, 4).copy(), ee[:4 * nee.value].reshape(-1, 4).copy(),
                    vg[:2 * nvg.value].reshape(-1, 2).copy())

    def toi(self, dq):
        t = C.c_double()
        dqa = np.ascontiguousarray(dq, dtype=np.float64).reshape(-1)
        self._check(self.lib.adb_step_size_upper_bound(self.h, dptr(dqa), C.byref(t)), "adb_step_size_upper_bound")
        return t.value

    def assemble(self, dense=True):
        n = 12 * self.B
        g = np.zeros(n)
        H = np.zeros(n * n) if dense else None
        self._check(self.lib.adb_assemble(self.h, dptr(g), dptr(H) if dense else None), "adb_assemble")
        return g, (H.reshape(n, n).T if dense else None)

    def contact_terms(self, q, dense=True):
        """Gradient and dense Hessian at the Newton iterate q of a step that
        starts at the current state (friction lagged from the contacts at q)."""
        n = 12 * self.B
        g = np.zeros(n)
        H = np.zeros(n * n) if dense else None
        qa = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        self._check(self.lib.adb_contact_terms(self.h, dptr(qa), dptr(g), dptr(H) if dense else None),
                    "adb_contact_terms")
        return g, (H.reshape(n, n).T if dense else None)

    def line_search(self, q, dq):
        """(status, alpha, E0, E1, trials) of the Armijo search from q along dq."""
        a, e0, e1, t = C.c_double(), C.c_double(), C.c_double(), C.c_int()
        qa = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        da = np.ascontiguousarray(dq, dtype=np.float64).reshape(-1)
        rc = self.lib.adb_line_search(self.h, dptr(qa), dptr(da), C.byref(a), C.byref(e0), C.byref(e1), C.byref(t))
        self._check(rc, "adb_line_search", allow=(-6,))
        return rc, a.value, e0.value, e1.value, t.value

    def adjoint_solve(self, rhs):
        """x = H^{-1} rhs with the exact Hessian of the last step's incremental
        potential at the current state (frictionless worlds)."""
        r = np.ascontiguousarray(rhs, dtype=np.float64).reshape(-1)
        x = np.zeros(12 * self.B)
        it = C.c_int()
        self._check(self.lib.adb_adjoint_solve(self.h, dptr(r), dptr(x), C.byref(it)), "adb_adjoint_solve")
        return x

    def profile_enable(self, on=True):
        self._check(self.lib.adb_profile_enable(self.h, 1 if on else 0), "adb_profile_enable")

    def profile_read(self):
        from .capi import PROF_KERNELS
        out = {}
        for k, name in enumerate(PROF_KERNELS):
            ms, n, b = C.c_double(), C.c_longlong(), C.c_double()
            self._check(self.lib.adb_profile_read(self.h, k, C.byref(ms), C.byref(n), C.byref(b)), "adb_profile_read")
            fl = C.c_double()
            self._check(self.lib.adb_profile_flops(self.h, k, C.byref(fl)), "adb_profile_flops")
            out[name] = {"ms": ms.value, "launches": n.value, "bytes": b.value, "flops": fl.value}
        return out

    def newton_direction(self):
        dq = np.zeros(12 * self.B)
        rc = self.lib.adb_newton_direction(self.h, dptr(dq))
        self._check(rc, "adb_newton_direction", allow=(-6,))
        return rc, dq
