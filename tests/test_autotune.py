"""The autotuner (adb_amd/autotune.py) on a stand-in world: coordinate
descent keeps the fastest candidate of each knob, restores the state it
started from, writes its choice to the cache and applies a cached choice
without timing.  (Its GPU test checks that every tuning gives bit-identical
steps.)"""
import numpy as np

from adb_amd import autotune


class FakeWorld:
    """Step time = a function of the tuning; the state advances per step."""

    def __init__(self):
        self.t = {"narrow_per_pair": 1.125, "ee_slices": 4, "pcg_batch": 16}
        self.q = np.zeros((2, 12))
        self.qd = np.zeros((2, 12))
        self.steps = 0

    def tuning(self):
        return dict(self.t)

    def set_tuning(self, **k):
        self.t.update(k)

    def state(self):
        return self.q.copy(), self.qd.copy()

    def set_state(self, q, qd=None):
        self.q = np.array(q)

    def step(self):
        import time
        cost = {0.5: 3, 1.125: 2, 2.0: 1}[self.t["narrow_per_pair"]] + {2: 1, 4: 3, 8: 2}[self.t["ee_slices"]] \
            + {8: 2, 16: 2, 32: 1}[self.t["pcg_batch"]]
        time.sleep(0.002 * cost)
        self.q += 1.0
        self.steps += 1
        return 0, None


def test_coordinate_descent_and_cache(tmp_path):
    cache = str(tmp_path / "tune.json")
    w = FakeWorld()
    t, src = autotune.autotune(w, key="k1", probe_steps=1, cache=cache)
    assert src == "timed"
    assert t == {"narrow_per_pair": 2.0, "ee_slices": 2, "pcg_batch": 32}
    assert np.all(w.q == 0.0)                      # state restored
    w2 = FakeWorld()
    t2, src2 = autotune.autotune(w2, key="k1", probe_steps=1, cache=cache)
    assert src2 == "cache" and t2 == t and w2.steps == 0
