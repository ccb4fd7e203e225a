"""Differentiable rollouts of the Newton contact step (hand-written adjoint).

Each step q_{n+1} = argmin_q E(q; q~_n) with q~_n = q_n + dt v_n + dt^2 g
(AffineBody::q_tile, model/mutils.cpp:245-252) and v_{n+1} = (q_{n+1} - q_n) / dt
(model/abd.cpp:128-134).  At the converged minimum grad E(q_{n+1}) = M (q_{n+1} - q~_n)
+ grad(dt^2 E_ortho + barrier)(q_{n+1}) = 0, so by implicit differentiation
dq_{n+1}/dq~_n = H^{-1} M with H the exact Hessian of E at q_{n+1}.  Given the
adjoints (l_q, l_v) of (q_{n+1}, v_{n+1}):

    a = l_q + l_v / dt,   w = H^{-1} a   (adb_adjoint_solve: same assembly and PCG as the step)
    l_q <- M w - l_v / dt,   l_v <- dt M w

Static bodies (mass < 0: q_{n+1} = q_n) have H = I on their rows and no
coupling, so with M = I there the same formulas give l_q <- l_q, and l_v <- 0.

The forward pass keeps a checkpoint (q, v) every `checkpoint_every` steps
in host memory; the backward pass recomputes each segment from its
checkpoint (the device path is deterministic, so the recomputed states are
bit-identical) and sweeps it in reverse.  Friction is lagged on the step's
start (model/friction.cpp), which this adjoint does not differentiate:
frictionless worlds only.
"""
import numpy as np
import torch

from .world import World


def mass_diag(scene):
    """Per body the diagonal of M (m x3, Ic x9); 1 on the rows of static bodies."""
    m = np.asarray(scene["mass"], dtype=np.float64)
    ic = np.asarray(scene["Ic"], dtype=np.float64)
    d = np.concatenate([np.repeat(m[:, None], 3, 1), np.repeat(ic[:, None], 9, 1)], axis=1)
    d[m < 0] = 1.0
    return d


class Rollout:
    """`steps` steps of a frictionless world from a given (q0, qdot0), with the
    vector-Jacobian product of the end state by checkpointed recomputation."""

    def __init__(self, scene, params, steps, checkpoint_every=10, world=None):
        """`world` (optional) is any object with the World methods step, state,
        set_state and adjoint_solve; by default a World on the GPU."""
        if params.mu != 0.0 and (params.pt_fric or params.ee_fric or params.vg_fric):
            raise ValueError("the adjoint covers frictionless steps only (mu = 0)")
        self.world = World(scene, params) if world is None else world
        self.steps = int(steps)
        self.every = max(1, int(checkpoint_every))
        self.dt = params.dt
        self.M = mass_diag(scene)
        self.static = np.asarray(scene["mass"]) < 0
        self.stats = []

    def _step(self):
        rc, st = self.world.step()
        if rc != 0:
            raise RuntimeError("step failed with status %d" % rc)
        self.stats.append(st.newton_iters)

    def forward(self, q0, qdot0):
        w = self.world
        w.set_state(q0, qdot0)
        ckpt = {0: (np.array(q0, dtype=np.float64).reshape(-1, 12), np.array(qdot0, dtype=np.float64).reshape(-1, 12))}
        for n in range(self.steps):
            self._step()
            if (n + 1) % self.every == 0 and n + 1 < self.steps:
                ckpt[n + 1] = w.state()
        q, qd = w.state()
        return q, qd, ckpt

    def backward(self, ckpt, lam_q, lam_v):
        w, dt = self.world, self.dt
        lq = np.array(lam_q, dtype=np.float64).reshape(-1, 12)
        lv = np.array(lam_v, dtype=np.float64).reshape(-1, 12)
        starts = sorted(ckpt)
        for s0 in reversed(starts):
            s1 = min(s0 + self.every, self.steps)
            w.set_state(*ckpt[s0])
            states = [ckpt[s0]]
            for _ in range(s0, s1):
                self._step()
                states.append(w.state())
            for n in range(s1 - 1, s0 - 1, -1):
                w.set_state(*states[n + 1 - s0])          # the Hessian at q_{n+1}
                a = lq + lv / dt
                x = w.adjoint_solve(a).reshape(-1, 12)
                Mx = self.M * x
                lq = Mx - lv / dt
                lv = dt * Mx
                lv[self.static] = 0.0
        return lq, lv


class _RolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q0, qdot0, sim):
        q, qd, ckpt = sim.forward(q0.detach().cpu().numpy(), qdot0.detach().cpu().numpy())
        ctx.sim, ctx.ckpt = sim, ckpt
        ctx.dev = q0.device
        return torch.from_numpy(q).to(q0.device), torch.from_numpy(qd).to(q0.device)

    @staticmethod
    def backward(ctx, gq, gv):
        # unused outputs arrive as zeros (gradients are materialized by default)
        lq, lv = ctx.sim.backward(ctx.ckpt, gq.detach().cpu().numpy(), gv.detach().cpu().numpy())
        return torch.from_numpy(lq).to(ctx.dev), torch.from_numpy(lv).to(ctx.dev), None


def rollout(q0, qdot0, sim):
    """(q_T, qdot_T) after sim.steps steps from (q0, qdot0) (float64 tensors,
    B x 12), differentiable with respect to both."""
    return _RolloutFn.apply(q0, qdot0, sim)


def centroid(scene, q, body):
    """World position of a body's rest-vertex centroid, x = A xbar_c + p
    (differentiable in q)."""
    vo = scene["vert_offset"]
    xc = torch.as_tensor(np.asarray(scene["rest_verts"][vo[body]:vo[body + 1]]).mean(0), dtype=q.dtype, device=q.device)
    qb = q[body]
    return qb[0:3] + qb[3:6] * xc[0] + qb[6:9] * xc[1] + qb[9:12] * xc[2]
