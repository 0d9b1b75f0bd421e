"""torch.autograd.Function over the library's differentiable rollout.

The forward runs ``substeps`` substeps of a ClothBody from (x0, v0) with the given
per-constraint stiffness (``cm_set_stiffness``, ``cm_forward_ckpt``, a checkpoint
every ``every`` substeps) and returns the final positions and velocities; the
backward calls ``cm_adjoint``, which recomputes each segment from its checkpoint and
sweeps it backwards on the GPU. Gradients flow to the distance and bend stiffness
and to the initial positions and velocities. Contact sets and Jacobi counts are
frozen per substep; the Morton / sort / LBVH / CCD stages carry no gradient.

Tensors cross the boundary through host memory (the C-ABI takes host arrays);
outputs come back on the device and in the dtype of ``x0``.

    roll = Rollout(body, substeps=60, every=10)
    x, v = roll(k_dist, k_bend, x0, v0)
    loss = ((x - target) ** 2).sum(dim=1).mean()
    loss.backward()
"""
from __future__ import annotations

import torch

from .body import ClothBody


def _host(t):
    return None if t is None else t.detach().to("cpu", torch.float32).contiguous().numpy()


class _RolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, k_dist, k_bend, x0, v0, body, substeps, every, tri0):
        body.set_stiffness(_host(k_dist), _host(k_bend))
        body.set_state(_host(x0), _host(v0), tri=tri0)
        body.forward_ckpt(substeps, every)
        x = torch.from_numpy(body.positions()).to(x0.device, x0.dtype)
        v = torch.from_numpy(body.velocities()).to(x0.device, x0.dtype)
        ctx.body = body
        ctx.meta = (x0.device, x0.dtype, k_dist.device, k_dist.dtype, k_bend.device,
                    k_bend.dtype)
        return x, v

    @staticmethod
    def backward(ctx, gx, gv):
        dev, dt, kd_dev, kd_dt, kb_dev, kb_dt = ctx.meta
        gx0, gv0, gkd, gkb = ctx.body.adjoint(_host(gx), _host(gv))
        return (torch.from_numpy(gkd).to(kd_dev, kd_dt), torch.from_numpy(gkb).to(kb_dev, kb_dt),
                torch.from_numpy(gx0).to(dev, dt), torch.from_numpy(gv0).to(dev, dt),
                None, None, None, None)


class Rollout:
    """Differentiable rollout of one body. The triangle order at construction is
    restored before every forward, so each call starts from the same state."""

    def __init__(self, body: ClothBody, substeps: int, every: int = 10):
        if substeps < 1 or every < 1:
            raise ValueError("substeps and every must be positive")
        self.body = body
        self.substeps = int(substeps)
        self.every = int(every)
        self.tri0 = body.buffer("tri").copy()

    def __call__(self, k_dist, k_bend, x0, v0):
        return _RolloutFn.apply(k_dist, k_bend, x0, v0, self.body, self.substeps, self.every,
                                self.tri0)
