"""Differentiable Poiseuille rollouts on the GPU.

    run = ChannelRollout(sc.params, sc.pos, sc.vel, steps=200, checkpoint_every=20)
    mu = torch.tensor(20.0, dtype=torch.float64, requires_grad=True)
    g = torch.tensor([1.8e-3, 0, 0], dtype=torch.float64, requires_grad=True)
    pos, vel = run(mu, g)
    loss = profile_loss(pos, vel, run.params, target)
    loss.backward()            # mu.grad, g.grad

The forward pass is the plain System's update() (HIP graphs between
checkpoints).  The backward pass is hand-written: the state is checkpointed
every `checkpoint_every` steps; each segment is recomputed forward from its
checkpoint and walked backwards with cmag_system_vjp_step, whose kernels
(adjoint.hip) are the vector-Jacobian products of density, force and
integrate with the step's neighbour structure held fixed.  The viscosity
enters as mu = rho0 * nu (the parameter block's mu), the body force as
gravity.  The time step stays the parameter block's delta_time.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from .systems import ParamSystem


def _f4(a, device) -> torch.Tensor:
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32).reshape(-1, 4), device=device).clone()


class ChannelRollout:
    """A Poiseuille system with a fixed initial state and step count; calling
    it with (mu, gravity) tensors returns the final (pos, vel) as a
    differentiable function of them (losses may depend on both; the type flag
    in pos[:, 3] carries no gradient)."""

    def __init__(self, params, pos, vel=None, vel_lf=None, steps: int = 200, checkpoint_every: int = 20,
                 use_graph: bool = True):
        if not torch.cuda.is_available():
            raise RuntimeError("ChannelRollout needs a GPU (the adjoint kernels are HIP only)")
        if params.system != L.SYS_POISEUILLE:
            raise ValueError("the adjoint covers the Poiseuille system")
        if steps < 1 or checkpoint_every < 1:
            raise ValueError("steps and checkpoint_every must be positive")
        self.params = params.copy()
        self.steps = int(steps)
        self.checkpoint_every = int(checkpoint_every)
        self.use_graph = use_graph
        self.n = int(np.asarray(pos).shape[0])
        dev = torch.device("cuda")
        self.device = dev
        vel = np.zeros_like(pos) if vel is None else vel
        vel_lf = vel if vel_lf is None else vel_lf
        self.state0 = (_f4(pos, dev), _f4(vel, dev), _f4(vel_lf, dev))
        self.system = ParamSystem(self.params, self.n)
        self._ncp = (self.steps + self.checkpoint_every - 1) // self.checkpoint_every
        # the recomputed states of one segment: scratch of a single backward
        # pass (autograd runs backward passes one after another)
        self._seg = [self._state() for _ in range(self.checkpoint_every)]

    def _state(self):
        return tuple(torch.empty((self.n, 4), dtype=torch.float32, device=self.device) for _ in range(3))

    def __call__(self, mu: torch.Tensor, gravity: torch.Tensor):
        return _Rollout.apply(mu, gravity, self)

    # state transfer between the System's arrays and torch tensors
    _ARRAYS = (L.ARRAY_POSITION, L.ARRAY_VELOCITY, L.ARRAY_VELOCITYLEAPFROG)

    def _put(self, state) -> None:
        for a, t in zip(self._ARRAYS, state):
            self.system.copy_from_device(a, t.data_ptr())

    def _take(self, state) -> None:
        for a, t in zip(self._ARRAYS, state):
            self.system.copy_to_device(a, t.data_ptr())

    def _params(self, mu: float, gravity) -> None:
        P = self.params.copy()
        P.mu = float(mu)
        for a in range(3):
            P.gravity[a] = float(gravity[a])
        self.system.set_params(P)


class _Rollout(torch.autograd.Function):
    """Each forward call keeps its own checkpoints on ctx, so several calls of
    one rollout (two parameter points in one loss) backpropagate through their
    own trajectories."""

    @staticmethod
    def forward(ctx, mu, gravity, run: ChannelRollout):
        g = gravity.detach().reshape(3).cpu().tolist()
        run._params(float(mu.detach().cpu()), g)
        s = run.system
        cps = [run._state() for _ in range(run._ncp)]
        torch.cuda.synchronize()
        run._put(run.state0)
        t, c = 0, 0
        while t < run.steps:
            run._take(cps[c])
            k = min(run.checkpoint_every, run.steps - t)
            s.run(k, use_graph=run.use_graph)
            t += k
            c += 1
        pos = torch.empty((run.n, 4), dtype=torch.float32, device=run.device)
        vel = torch.empty_like(pos)
        s.copy_to_device(L.ARRAY_POSITION, pos.data_ptr())
        s.copy_to_device(L.ARRAY_VELOCITY, vel.data_ptr())
        s.synchronize()
        ctx.run = run
        ctx.cps = cps
        ctx.theta = (float(mu.detach().cpu()), g)
        ctx.mu_like = (mu.dtype, mu.device, mu.shape)
        ctx.g_like = (gravity.dtype, gravity.device, gravity.shape)
        return pos, vel

    @staticmethod
    def backward(ctx, gpos_out, gvel_out):
        run: ChannelRollout = ctx.run
        run._params(*ctx.theta)
        s = run.system
        n, dev = run.n, run.device

        def seed(g):
            # the adjoint of the final state; w (the type flag) is not a variable
            out = torch.zeros((n, 4), dtype=torch.float32, device=dev)
            if g is not None:
                out[:, :3] = g.detach()[:, :3].to(torch.float32)
            return out

        gpos = seed(gpos_out)
        gvel = seed(gvel_out)
        glf = torch.zeros_like(gpos)
        gmu = torch.zeros(n, dtype=torch.float64, device=dev)
        ggrav = torch.zeros((n, 4), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        ce = run.checkpoint_every
        for c in reversed(range(len(ctx.cps))):
            t0 = c * ce
            t1 = min(t0 + ce, run.steps)
            run._put(ctx.cps[c])
            for t in range(t0, t1):
                run._take(run._seg[t - t0])
                if t < t1 - 1:
                    s.update()
            for t in reversed(range(t0, t1)):
                run._put(run._seg[t - t0])
                s.vjp_step(gpos.data_ptr(), gvel.data_ptr(), glf.data_ptr(), gmu.data_ptr(), ggrav.data_ptr())
        s.synchronize()
        dtype, device, shape = ctx.mu_like
        dmu = gmu.sum().to(dtype=dtype, device=device).reshape(shape)
        dtype, device, shape = ctx.g_like
        dg = ggrav[:, :3].sum(dim=0).to(dtype=dtype, device=device).reshape(shape)
        return dmu, dg, None


def profile_loss(pos: torch.Tensor, vel: torch.Tensor, params, target, bins: int = 64) -> torch.Tensor:
    """sum over non-empty y-bins of (mean x-velocity of the fluid in the bin -
    target)^2, the bins splitting [wall_bottom, wall_top] evenly"""
    fl = pos[:, 3] == 0.0
    lo, hi = float(params.wall_bottom), float(params.wall_top)
    y = pos[fl, 1].to(torch.float64)
    b = torch.clamp(torch.floor((y - lo) / ((hi - lo) / bins)).to(torch.int64), 0, bins - 1)
    u = vel[fl, 0].to(torch.float64)
    cnt = torch.zeros(bins, dtype=torch.float64, device=pos.device).index_add_(0, b, torch.ones_like(u))
    s = torch.zeros(bins, dtype=torch.float64, device=pos.device).index_add(0, b, u)
    nz = cnt > 0
    tgt = torch.as_tensor(np.asarray(target, np.float64), device=pos.device)
    prof = s[nz] / cnt[nz]
    return ((prof - tgt[nz]) ** 2).sum()


def profile(pos: torch.Tensor, vel: torch.Tensor, params, bins: int = 64) -> np.ndarray:
    """the binned mean x-velocity profile (nan in empty bins)"""
    fl = (pos[:, 3] == 0.0).cpu().numpy()
    p = pos.detach().cpu().numpy()[fl]
    v = vel.detach().cpu().numpy()[fl]
    lo, hi = float(params.wall_bottom), float(params.wall_top)
    b = np.clip(np.floor((p[:, 1].astype(np.float64) - lo) / ((hi - lo) / bins)).astype(np.int64), 0, bins - 1)
    cnt = np.bincount(b, minlength=bins).astype(np.float64)
    s = np.bincount(b, weights=v[:, 0].astype(np.float64), minlength=bins)
    out = np.full(bins, np.nan)
    out[cnt > 0] = s[cnt > 0] / cnt[cnt > 0]
    return out


__all__ = ["ChannelRollout", "profile_loss", "profile"]
