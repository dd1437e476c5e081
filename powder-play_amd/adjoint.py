"""Differentiable rollout of the fixed-corotated elastic solid.

``ElasticRollout`` runs ``steps`` MLS-MPM steps of a solid whose stress is the
reference's snow stress at hardening factor 1 without plastic projection
(src/lib/mpm/reference/solver.ts:2171-2203) — fixed-corotated elasticity — on
the GPU, keeps a checkpoint every ``segment_steps`` steps, and runs the
hand-written reverse pass through ``mpm_diff_backward``.  ``rollout`` wraps the
pair as a ``torch.autograd.Function``: gradients reach the initial state and
the per-particle lambda and mu.

State rows are 12 floats: x, y, vx, vy, C00, C01, C10, C11, F00, F01, F10, F11.
NumPy inputs are copied from host memory; CUDA tensors are read in place on
the device.  The reference has no differentiation code: this module is an
addition whose forward step follows the solid branch of ``stepMpm``
(solver.ts:396-824).
"""
from __future__ import annotations

import numpy as np

from . import native

STATE = native.MPM_DIFF_STATE_FLOATS


def _torch():
    import torch
    return torch


def _is_tensor(a) -> bool:
    try:
        torch = _torch()
    except ImportError:
        return False
    return isinstance(a, torch.Tensor)


class ElasticRollout:
    """GPU forward and reverse of a fixed-corotated rollout (one config)."""

    def __init__(self, particle_count: int, config: dict, segment_steps: int = 20, device: int = -1,
                 graphs: bool = True):
        g = config.get("gravity", (0.0, 0.0))
        cfg = native.MpmDiffConfig(
            particle_count, config["cellsX"], config["cellsY"], config["steps"], segment_steps,
            config["cellSize"], config["dt"], config["particleMass"], config["particleVolume"],
            float(g[0]), float(g[1]), config.get("wallFriction", 0.0), device,
            0 if graphs else native.MPM_DIFF_NO_GRAPHS)
        self.n = int(particle_count)
        self.steps = int(config["steps"])
        self.native = native.DiffRollout(cfg)

    # -- argument plumbing: host arrays or CUDA tensors, float32, contiguous
    def _in(self, a, shape):
        if _is_tensor(a):
            torch = _torch()
            t = a.detach().to(torch.float32).contiguous()
            if tuple(t.shape) != shape:
                raise ValueError(f"expected shape {shape}, got {tuple(t.shape)}")
            if t.is_cuda:
                torch.cuda.current_stream(t.device).synchronize()
                return t, t.data_ptr()
            t = t.numpy()
            return t, t.ctypes.data
        arr = np.ascontiguousarray(a, np.float32)
        if arr.shape != shape:
            raise ValueError(f"expected shape {shape}, got {arr.shape}")
        return arr, arr.ctypes.data

    def _out(self, like, shape):
        if _is_tensor(like) and like.is_cuda:
            torch = _torch()
            t = torch.empty(shape, dtype=torch.float32, device=like.device)
            return t, t.data_ptr()
        arr = np.empty(shape, np.float32)
        return arr, arr.ctypes.data

    def forward(self, state0, lam, mu):
        """Final state (n, 12) of the rollout from state0 with per-particle
        lambda and mu (n,)."""
        s, sp = self._in(state0, (self.n, STATE))
        la, lp = self._in(lam, (self.n,))
        m, mp = self._in(mu, (self.n,))
        out, op = self._out(state0, (self.n, STATE))
        self.native.forward(sp, lp, mp, op)
        del s, la, m
        return out

    def backward(self, bar_final):
        """(dL/dstate0 (n, 12), dL/dlambda (n,), dL/dmu (n,)) for
        dL/d(final state) = bar_final, through the last forward."""
        b, bp = self._in(bar_final, (self.n, STATE))
        b0, b0p = self._out(bar_final, (self.n, STATE))
        gl, glp = self._out(bar_final, (self.n,))
        gm, gmp = self._out(bar_final, (self.n,))
        self.native.backward(bp, b0p, glp, gmp)
        del b
        return b0, gl, gm

    def status(self) -> int:
        return self.native.status()

    def timing(self) -> dict:
        f, b = self.native.timing()
        return {"forward_ms": f, "backward_ms": b}

    def layout(self) -> dict:
        seg, ck, nbytes = self.native.layout()
        return {"segment_steps": seg, "checkpoints": ck, "device_bytes": nbytes}

    def close(self):
        self.native.close()


def _make_function():
    torch = _torch()

    class RolloutFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, runner, state0, lam, mu):
            ctx.runner = runner
            return runner.forward(state0, lam, mu)

        @staticmethod
        def backward(ctx, grad_final):
            b0, gl, gm = ctx.runner.backward(grad_final)
            return None, b0, gl, gm

    return RolloutFunction


_FUNCTION = None


def rollout(runner: ElasticRollout, state0, lam, mu):
    """Autograd form: final state as a tensor whose backward runs the GPU
    reverse pass.  A second forward through the same runner replaces the
    checkpoints the pending backward would use, so call backward first."""
    global _FUNCTION
    if _FUNCTION is None:
        _FUNCTION = _make_function()
    return _FUNCTION.apply(runner, state0, lam, mu)


def com_loss(final_state, target):
    """|mean(x) - target|^2 of equal-mass particles (NumPy or torch)."""
    if _is_tensor(final_state):
        torch = _torch()
        t = torch.as_tensor(target, dtype=final_state.dtype, device=final_state.device)
        d = final_state[:, :2].mean(dim=0) - t
        return (d * d).sum()
    d = np.asarray(final_state, np.float64)[:, :2].mean(axis=0) - np.asarray(target, np.float64)
    return float(d @ d)
