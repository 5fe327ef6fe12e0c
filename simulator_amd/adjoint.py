"""Differentiable cloth rollouts over the C ABI (cc_adj_*).

`Adjoint` runs a rollout of the fused step and its hand-written adjoint on
the GPU; `ClothRollout` wraps it as a torch.autograd.Function:

    adj = Adjoint(sim, nsteps=200, ckpt_every=20)
    x = rollout(adj, theta, v0)             # theta = (k0, k1, k2, damp, rest_scale)
    loss = ((x - target) ** 2).mean()
    loss.backward()                         # theta.grad (float64), v0.grad

The gradient holds the forward's contacts fixed (each projected vertex keeps
its plane projection; the contact normals entering the force projection are
constants), see include/clothcore.h.
"""
import ctypes as C

import numpy as np

from . import _lib as L


class Adjoint:
    """A rollout of `nsteps` frames over `sim`'s cloth, springs, pins and
    (static) body, with its adjoint; states checkpointed every
    `ckpt_every` steps."""

    def __init__(self, sim, nsteps, ckpt_every=20):
        self.sim = sim                       # keeps the simulator alive
        self.n = sim.n
        self.nsteps = int(nsteps)
        h = C.c_void_p()
        L.check(L.lib().cc_adj_create(sim._h, self.nsteps, int(ckpt_every), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            L.lib().cc_adj_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, x0, v0, theta):
        """Final positions (n, 3) of the rollout from x0 with initial
        velocity v0 (or None) and parameters theta."""
        x0 = np.ascontiguousarray(x0, np.float32)
        v0 = None if v0 is None else np.ascontiguousarray(v0, np.float32)
        th = np.ascontiguousarray(theta, np.float64)
        assert x0.shape == (self.n, 3) and th.shape == (5,)
        xf = np.zeros((self.n, 3), np.float32)
        L.check(L.lib().cc_adj_forward(self._h, L.ptr(x0), None if v0 is None else L.ptr(v0), L.ptr(th),
                                       L.ptr(xf)))
        return xf

    def backward(self, dl_dx):
        """(dL/dtheta float64 (5,), dL/dv0 (n, 3), dL/dx0 (n, 3)) of the
        last forward, given dL/dx_final."""
        g = np.ascontiguousarray(dl_dx, np.float32)
        assert g.shape == (self.n, 3)
        dth = np.zeros(5, np.float64)
        dv0 = np.zeros((self.n, 3), np.float32)
        dx0 = np.zeros((self.n, 3), np.float32)
        L.check(L.lib().cc_adj_backward(self._h, L.ptr(g), L.ptr(dth), L.ptr(dv0), L.ptr(dx0)))
        return dth, dv0, dx0


def _function():
    import torch

    class ClothRollout(torch.autograd.Function):
        """x_final = rollout(theta, v0) through the GPU forward and adjoint."""

        @staticmethod
        def forward(ctx, theta, v0, adj, x0):
            xf = adj.forward(x0, v0.detach().cpu().numpy(), theta.detach().cpu().numpy())
            ctx.adj = adj
            ctx.devices = (theta.device, v0.device)
            return torch.from_numpy(xf).to(v0.device)

        @staticmethod
        def backward(ctx, grad_x):
            dth, dv0, _ = ctx.adj.backward(grad_x.detach().cpu().numpy())
            return (torch.from_numpy(dth).to(ctx.devices[0]), torch.from_numpy(dv0).to(ctx.devices[1]),
                    None, None)

    return ClothRollout


_cached = None


def ClothRollout():
    """The torch.autograd.Function class (torch is imported on first use)."""
    global _cached
    if _cached is None:
        _cached = _function()
    return _cached


def rollout(adj, theta, v0, x0=None):
    """Differentiable final positions: theta (5,) float64 tensor, v0 (n, 3)
    float32 tensor; x0 defaults to the simulator's initial cloth."""
    if x0 is None:
        x0 = adj.sim._v4[:, :3]
    return ClothRollout().apply(theta, v0, adj, np.ascontiguousarray(x0, np.float32))
