"""Differentiable PBF rollouts (gradient work only).

:class:`Rollout` wraps the C ABI ``pbf_adj_*`` of include/pbf.h: a forward
pass of ``nsteps`` steps that keeps a checkpoint every ``checkpoint_every``
steps, and a backward pass that recomputes each segment from its checkpoint
(neighbour lists rebuilt) and runs the hand-written adjoint kernels.

:class:`PBFRollout` is the ``torch.autograd.Function`` over it: inputs are
the initial velocities (n, 3) and the tensile coefficient k (scalar), with
the initial positions held fixed; outputs are the final positions and
velocities.  Torch only carries tensors here; all arithmetic is in the HIP
library.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import PbfError, lib
from .params import Params


def _f(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Rollout:
    def __init__(self, params: Params, n: int, nsteps: int, checkpoint_every: int = 10):
        self.params = params
        self.n = int(n)
        self.nsteps = int(nsteps)
        self._h = C.c_void_p()
        # Number of forward passes run so far; the autograd binding records it
        # so that a backward pass is only taken against the forward it follows.
        self.forward_count = 0
        cp = params.to_c()
        self._check(lib().pbf_adj_create(C.byref(cp), self.n, self.nsteps, int(checkpoint_every),
                                         C.byref(self._h)))

    def _check(self, rc: int) -> None:
        if rc < 0:
            msg = lib().pbf_adj_last_error(self._h if self._h else None)
            raise PbfError(rc, msg.decode() if msg else "")

    def _rows(self, a, name: str) -> np.ndarray:
        """a as a contiguous float32 (n, 3) array; any other shape is refused
        (the library reads and writes n * 3 floats)."""
        a = np.ascontiguousarray(a, np.float32)
        if a.shape != (self.n, 3):
            raise ValueError(f"{name} has shape {a.shape}, expected ({self.n}, 3)")
        return a

    def forward(self, x0, v0, k: float):
        x0 = self._rows(x0, "x0")
        v0 = self._rows(v0, "v0")
        xT = np.empty_like(x0)
        vT = np.empty_like(v0)
        self._check(lib().pbf_adj_forward(self._h, _f(x0), _f(v0), float(k), _f(xT), _f(vT)))
        self.forward_count += 1
        return xT, vT

    def backward(self, xbar_T, vbar_T=None):
        """Gradients of the last forward pass."""
        xb = self._rows(xbar_T, "xbar_T")
        vb = None if vbar_T is None else self._rows(vbar_T, "vbar_T")
        xb0 = np.empty_like(xb)
        vb0 = np.empty_like(xb)
        kb = C.c_double()
        self._check(lib().pbf_adj_backward(self._h, _f(xb), None if vb is None else _f(vb),
                                           _f(xb0), _f(vb0), C.byref(kb)))
        return xb0, vb0, kb.value

    def stats(self):
        ms, contacts = C.c_double(), C.c_int64()
        self._check(lib().pbf_adj_stats(self._h, C.byref(ms), C.byref(contacts)))
        return ms.value, contacts.value

    def replay_mismatches(self) -> int:
        """Particle checks of the last backward pass whose replayed state
        differed from the forward pass at a checkpoint (0: the tape the
        adjoint ran over is the forward trajectory, bit for bit)."""
        m = C.c_int64()
        self._check(lib().pbf_adj_replay_check(self._h, C.byref(m)))
        return m.value

    def close(self) -> None:
        if self._h:
            lib().pbf_adj_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rollout_function():
    """Returns the torch.autograd.Function class (torch imported lazily)."""
    import torch

    class PBFRollout(torch.autograd.Function):
        @staticmethod
        def forward(ctx, v0, k, x0, rollout):
            xT, vT = rollout.forward(x0.detach().cpu().numpy(), v0.detach().cpu().numpy(),
                                     float(k.detach().cpu()))
            ctx.rollout = rollout
            ctx.forward_id = rollout.forward_count
            dev = v0.device
            return torch.from_numpy(xT).to(dev), torch.from_numpy(vT).to(dev)

        @staticmethod
        def backward(ctx, gx, gv):
            # The rollout's tape holds its last forward pass only.
            if ctx.rollout.forward_count != ctx.forward_id:
                raise RuntimeError("the Rollout ran another forward pass after this one; "
                                   "run backward before the next forward, or use one Rollout "
                                   "per graph")
            gxn = gx.detach().cpu().numpy() if gx is not None else None
            gvn = gv.detach().cpu().numpy() if gv is not None else None
            if gxn is None:
                gxn = np.zeros((ctx.rollout.n, 3), np.float32)
            xb0, vb0, kb = ctx.rollout.backward(gxn, gvn)
            dev = gx.device if gx is not None else gv.device
            return (torch.from_numpy(vb0).to(dev), torch.tensor(kb, dtype=torch.float32, device=dev),
                    torch.from_numpy(xb0).to(dev), None)

    return PBFRollout
