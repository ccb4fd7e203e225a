"""Autotuner for the launch shapes of a world (adb_tuning in include/adb_capi.h).

The knobs change how fast a step runs, never its results: the narrow phase's
workgroups per overlapping body pair (`narrow_per_pair`), the slices of its
edge-edge pass (`ee_slices`) and the PCG iterations queued between host
checks (`pcg_batch`).  The tuner times candidate values on the workload's own
state: it saves the world's state, and for each knob in turn (coordinate
descent, the others held at their current best) restores the state, applies a
candidate, runs `probe_steps` steps and keeps the fastest; the world's state
is restored at the end.  The choice is cached per workload (body, vertex,
face and edge counts, the parameters that shape the work, the device) in a
JSON file, so later runs of the same workload apply it without timing.
"""
import hashlib
import json
import os
import time

CANDIDATES = {
    "narrow_per_pair": [0.5, 1.125, 2.0],
    "ee_slices": [2, 4, 8],
    "pcg_batch": [8, 16, 32],
}


def default_cache_path():
    return os.environ.get("ADB_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "adb_amd", "tuning.json"))


def workload_key(scene, params, device=""):
    """Key of a workload: its sizes and the parameters that change the work."""
    import numpy as np
    parts = [len(scene["mass"]), int(scene["vert_offset"][-1]), int(scene["face_offset"][-1]),
             int(scene["edge_offset"][-1]), params.d_hat, params.dt, params.broad_phase_fp32, params.iaabb,
             params.enable_pt, params.enable_ee, params.enable_ground, device]
    h = hashlib.sha1(repr(parts).encode() + np.ascontiguousarray(scene["q0"]).tobytes()[:4096]).hexdigest()[:16]
    return "B%d-V%d-%s" % (parts[0], parts[1], h)


def _load(path):
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store(path, data):
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp%d" % os.getpid()
    with open(tmp, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def autotune(world, key=None, probe_steps=2, candidates=None, cache=None, log=None):
    """Choose and apply the tuning of `world` (an adb_amd.world.World or any
    object with step, state, set_state, tuning and set_tuning).  Returns
    (tuning dict, source) with source "cache" or "timed"."""
    cache = cache or default_cache_path()
    if key is not None:
        hit = _load(cache).get(key)
        if hit is not None:
            world.set_tuning(**hit)
            return world.tuning(), "cache"
    cand = candidates or CANDIDATES
    q, qd = world.state()
    best = dict(world.tuning())

    def run(t):
        world.set_state(q, qd)
        world.set_tuning(**t)
        t0 = time.perf_counter()
        for _ in range(probe_steps):
            world.step()
        return time.perf_counter() - t0

    run(best)                          # first-call set-up (buffers, captured graphs) outside the timings
    best_t = run(best)
    for knob, values in cand.items():
        for v in values:
            if v == best[knob]:
                continue
            t = dict(best, **{knob: v})
            el = run(t)
            if log:
                log("tune %s=%s: %.2f ms / step" % (knob, v, 1e3 * el / probe_steps))
            if el < best_t:
                best, best_t = t, el
    world.set_state(q, qd)
    world.set_tuning(**best)
    if key is not None:
        data = _load(cache)
        data[key] = best
        _store(cache, data)
    return world.tuning(), "timed"
