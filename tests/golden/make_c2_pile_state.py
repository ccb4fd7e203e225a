"""Writes tests/golden/c2_pile_step70.npz: the state (q, qdot; 2000 x 12 each)
of the 2000-icosphere workload (subdivision 3, r ~ U(0.05, 0.15), 4 x 4 x 8
box, ground, mu = 0.3, dt = 0.01, d_hat = 1e-3, seed 2; mixed precision) after
70 steps, when the spheres have piled up on the ground (~2.4M active
contacts, 4-point contacts and 96 x 96 preconditioner groups present).

It is an INPUT fixture for the parity tests at a late, dense state: the
expected outputs are computed by the oracle at test time from this state.
Run once on one MI355X from the repository root:

    python tests/golden/make_c2_pile_state.py [steps] [output.npz]
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)


def main():
    steps = int(sys.argv[1]) if len(sys.argv) > 1 else 70
    import bench
    from adb_amd.world import World
    sc = bench.workload(2000)
    w = World(sc, bench.params("mixed"))
    for s in range(steps):
        rc, st = w.step()
        if (s + 1) % 10 == 0:
            print("step %d: %d contacts, %d Newton iterations" % (s + 1, st.n_pt + st.n_ee + st.n_vg, st.newton_iters),
                  flush=True)
    q, qd = w.state()
    out = sys.argv[2] if len(sys.argv) > 2 else os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                              "c2_pile_step%d.npz" % steps)
    np.savez_compressed(out, q=q, qdot=qd, steps=steps)
    print("wrote", out)


if __name__ == "__main__":
    main()
