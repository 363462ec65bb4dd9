"""Differentiable PHOENIX time evolution: torch.autograd.Function over the HIP adjoint.

    psi_T, n_T = rk4_evolve(P, V, psi0, n0, physics=ph, dt=0.04, steps=2000, checkpoint_every=50)
    J = ((psi_T.abs() ** 2) * target).sum()
    J.backward()            # P.grad, V.grad (and psi0.grad, n0.grad if they require grad)

The forward pass runs `steps` fixed RK4 steps of the scalar GP + reservoir model
(RungeKutta4::step of the reference, runge_kutta.cu:74-94) on the GPU and keeps a
checkpoint every `checkpoint_every` steps; the backward pass recomputes the states
of one checkpoint interval at a time and runs the hand-written discrete-adjoint
kernel (phx_adjoint.hip) backwards through them.  psi0 / n0 may carry a leading
batch dimension (e.g. noise seeds); P and V are then shared by the batch and their
gradients summed over it, unless they carry the same batch dimension.

Complex gradients follow torch's convention, dJ/dRe + i dJ/dIm.  There is no CPU
path: tensors must live on the HIP device.
"""
from __future__ import annotations

from typing import Optional, Tuple

import torch

from . import _lib


class _Engine:
    """Handle cache keyed by the configuration (allocation of the checkpoint store is the expensive part)."""

    _cache: dict = {}

    @classmethod
    def get(cls, key, make):
        e = cls._cache.get(key)
        if e is None:
            e = make()
            cls._cache.clear()  # one live store at a time: checkpoints of large batches are GBs
            cls._cache[key] = e
        return e


def _physics_key(ph: _lib.PhxPhysics):
    return tuple(getattr(ph, k) for k, _ in _lib.PhxPhysics._fields_)


class _EngineStream:
    """Run the engine on torch's current stream, or, when that is the default stream (handle 0, which
    the handles' set_stream cannot take), on a side stream ordered after the current stream's work and
    before its later work by events; tensors the engine touches are recorded on the side stream so the
    caching allocator does not hand their memory out while the engine may still use it."""

    _side = None

    def __init__(self, engine, *tensors):
        self.engine, self.tensors = engine, tensors

    def __enter__(self):
        self.cur = torch.cuda.current_stream()
        if self.cur.cuda_stream != 0:
            self.engine.set_stream(self.cur.cuda_stream)
            self.side = None
            return self
        if _EngineStream._side is None:
            _EngineStream._side = torch.cuda.Stream()
        self.side = _EngineStream._side
        self.side.wait_stream(self.cur)
        self.engine.set_stream(self.side.cuda_stream)
        return self

    def __exit__(self, *exc):
        if self.side is not None:
            for t in self.tensors:
                if t is not None:
                    t.record_stream(self.side)
            self.cur.wait_stream(self.side)
        return False


class RK4Adjoint(torch.autograd.Function):
    """(P, V, psi0, n0) -> (psi_T, n_T) with a hand-written adjoint backward."""

    @staticmethod
    def forward(ctx, pump, pot, psi0, n0, engine, periodic):  # noqa: D401
        for t in (pump, pot, psi0, n0):
            if not t.is_cuda or not t.is_contiguous():
                raise ValueError("rk4_evolve expects contiguous tensors on the HIP device")
        psi_T = torch.empty_like(psi0)
        n_T = torch.empty_like(n0)
        with _EngineStream(engine, pump, pot, psi0, n0, psi_T, n_T):
            engine.forward(psi0.data_ptr(), n0.data_ptr(), pump.data_ptr(), pot.data_ptr(), psi_T.data_ptr(), n_T.data_ptr())
        engine.generation = getattr(engine, "generation", 0) + 1
        ctx.engine = engine
        ctx.generation = engine.generation
        ctx.save_for_backward(pump, pot, psi0, n0)
        return psi_T, n_T

    @staticmethod
    def backward(ctx, g_psi_T, g_n_T):
        pump, pot, psi0, n0 = ctx.saved_tensors  # also checks that P, V were not modified in place
        if g_psi_T is None:
            g_psi_T = torch.zeros_like(psi0)
        g_psi_T = g_psi_T.contiguous()
        g_n_ptr = None
        if g_n_T is not None:
            g_n_T = g_n_T.contiguous()
            g_n_ptr = g_n_T.data_ptr()
        gP = torch.empty_like(pump) if ctx.needs_input_grad[0] else None
        gV = torch.empty_like(pot) if ctx.needs_input_grad[1] else None
        gpsi = torch.empty_like(psi0) if ctx.needs_input_grad[2] else None
        gn = torch.empty_like(n0) if ctx.needs_input_grad[3] else None
        e = ctx.engine
        scratch_psi, scratch_n = torch.empty_like(psi0), torch.empty_like(n0)
        with _EngineStream(e, pump, pot, psi0, n0, g_psi_T, g_n_T, gP, gV, gpsi, gn, scratch_psi, scratch_n):
            if e.generation != ctx.generation:
                # another forward pass reused the checkpoint store since: rebuild this pass's checkpoints
                e.forward(psi0.data_ptr(), n0.data_ptr(), pump.data_ptr(), pot.data_ptr(), scratch_psi.data_ptr(),
                          scratch_n.data_ptr())
                e.generation += 1
                ctx.generation = e.generation
            e.backward(g_psi_T.data_ptr(), g_n_ptr, gP.data_ptr() if gP is not None else None,
                       gV.data_ptr() if gV is not None else None, gpsi.data_ptr() if gpsi is not None else None,
                       gn.data_ptr() if gn is not None else None)
        return gP, gV, gpsi, gn, None, None


def rk4_evolve(pump: torch.Tensor, pot: torch.Tensor, psi0: torch.Tensor, n0: torch.Tensor, *, physics: _lib.PhxPhysics,
               dt: float, steps: int, checkpoint_every: int = 50, periodic: Tuple[bool, bool] = (True, True)
               ) -> Tuple[torch.Tensor, torch.Tensor]:
    """Differentiable `steps` RK4 steps.  psi0: complex (N_r, N_c) or (B, N_r, N_c); n0, P, V real.

    FP32 for complex64 / float32 inputs, FP64 for complex128 / float64."""
    if psi0.dtype not in (torch.complex64, torch.complex128):
        raise TypeError("psi0 must be complex64 or complex128")
    real = torch.float32 if psi0.dtype == torch.complex64 else torch.float64
    for name, t in (("n0", n0), ("pump", pump), ("pot", pot)):
        if t.dtype != real:
            raise TypeError(f"{name} must be {real}")
    batched = psi0.dim() == 3
    B = psi0.shape[0] if batched else 1
    N_r, N_c = psi0.shape[-2:]
    if n0.shape != psi0.shape:
        raise ValueError("n0 must have the shape of psi0")
    shared = pump.dim() == 2
    if pump.shape != pot.shape or pump.shape[-2:] != (N_r, N_c) or (not shared and pump.shape[0] != B):
        raise ValueError("pump and pot must be (N_r, N_c) or (B, N_r, N_c)")
    prec = "fp32" if real == torch.float32 else "fp64"
    key = (prec, _physics_key(physics), N_c, N_r, B, float(dt), int(steps), int(checkpoint_every), tuple(periodic), shared)
    engine = _Engine.get(key, lambda: _lib.Adjoint(prec, physics, N_c, N_r, B, dt, steps, checkpoint_every, periodic,
                                                   reservoir=True, pump=True, potential=True, shared_params=shared))
    return RK4Adjoint.apply(pump.contiguous(), pot.contiguous(), psi0.contiguous(), n0.contiguous(), engine, tuple(periodic))


def release_engines():
    """Free the cached checkpoint store."""
    for e in _Engine._cache.values():
        e.close()
    _Engine._cache.clear()
