"""Differentiable leapfrog: gradients of a direct-sum N-body run with respect to the
initial positions, velocities and masses, through a hand-written adjoint.

    x, v = leapfrog(x0, v0, m, dt, eps, nsteps, checkpoint_every=20)
    loss = ((x - target) ** 2).sum(-1).mean(); loss.backward()

x0, v0: (n,3) float64 CUDA tensors (kpc, km/s); m: (n,) float64 (Msun); dt in Myr,
eps in kpc.  The forward is the drift-kick-drift of Simulation.run (gh_adjoint_forward,
the same kernels as gh_run) and keeps (x, v) every `checkpoint_every` steps; backward
re-runs each segment from its checkpoint and applies the reverse step with the
vector-Jacobian products through the softened tidal tensor (gh_adjoint_backward).
Both are ordered with torch's current stream.
"""
import ctypes
import math

import torch

from . import _native


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


class _OnTorchStream:
    """Orders library work with torch's current stream.  A real stream is handed to the
    library; torch's legacy default stream (handle 0) cannot be, since 0 selects the
    context's own stream, so then the call is fenced by synchronisation on both sides."""

    def __init__(self, lib, device):
        self.lib = lib
        self.torch_stream = torch.cuda.current_stream(device)
        self.handle = self.torch_stream.cuda_stream

    def __enter__(self):
        if self.handle == 0:
            self.torch_stream.synchronize()
        return ctypes.c_void_p(self.handle or None)

    def __exit__(self, *exc):
        if self.handle == 0:
            self.lib.sync()
        return False


class DirectLeapfrog(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, m, dt, eps, nsteps, every, precision):
        lib = _native.default_context()
        n = x0.shape[0]
        x = x0.detach().to(torch.float64).contiguous().clone()
        v = v0.detach().to(torch.float64).contiguous().clone()
        mm = m.detach().to(torch.float64).contiguous()
        nseg = max(1, math.ceil(nsteps / every))
        ckpt = torch.empty((nseg, 2, n, 3), dtype=torch.float64, device=x.device)
        with _OnTorchStream(lib, x.device) as stream:
            lib.check(lib.lib.gh_adjoint_forward(lib.handle, _p(x), _p(v), _p(mm), n, float(dt),
                                                 float(eps), int(precision), int(nsteps),
                                                 int(every), _p(ckpt), stream))
        ctx.save_for_backward(mm, ckpt)
        ctx.params = (float(dt), float(eps), int(precision), int(nsteps), int(every))
        return x, v

    @staticmethod
    def backward(ctx, gx, gv):
        mm, ckpt = ctx.saved_tensors
        dt, eps, precision, nsteps, every = ctx.params
        n = mm.shape[0]
        dev = mm.device
        lx = (gx if gx is not None else torch.zeros((n, 3), device=dev)).to(
            torch.float64).contiguous().clone()
        lv = (gv if gv is not None else torch.zeros((n, 3), device=dev)).to(
            torch.float64).contiguous().clone()
        lm = torch.empty(n, dtype=torch.float64, device=dev)
        lib = _native.default_context()
        with _OnTorchStream(lib, dev) as stream:
            lib.check(lib.lib.gh_adjoint_backward(lib.handle, _p(ckpt), _p(mm), n, dt, eps,
                                                  precision, nsteps, every, _p(lx), _p(lv),
                                                  _p(lm), stream))
        return lx, lv, lm, None, None, None, None, None


def leapfrog(x0, v0, m, dt, eps, nsteps, checkpoint_every=20, precision="fp32"):
    """Final (x, v) after nsteps drift-kick-drift steps, differentiable in x0, v0, m."""
    return DirectLeapfrog.apply(x0, v0, m, dt, eps, int(nsteps), int(checkpoint_every),
                                _native.PRECISIONS[precision])
