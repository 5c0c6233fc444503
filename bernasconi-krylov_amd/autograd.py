"""Differentiable Trotter evolution: a torch.autograd.Function over the adjoint
entry points of libbk (bk_adjoint_forward / bk_adjoint_backward).

The field is piecewise constant, g(t) = g_j on segment j (steps_per_segment
Trotter steps each).  Forward returns the tensor [P_<(T), <d_H>(T)]; backward
turns the incoming gradient (w_p, w_d) into the objective weights of one
reverse pass, so any linear combination of the two observables costs one
adjoint sweep.  The reference has no differentiable path; this is new.

PyTorch is used on the Python side only: the g_j live in a float64 tensor
(any device), the state never leaves libbk's device buffers.
"""
from __future__ import annotations

import importlib
from typing import Optional

import numpy as np
import torch

_bk = importlib.import_module("bernasconi-krylov_amd")


class TrotterObservables(torch.autograd.Function):
    """[P_<(T), <d_H>(T)] as a function of the segment fields g (1-D tensor)."""

    @staticmethod
    def forward(ctx, g_segments: torch.Tensor, engine, n: int, start_bits: int, dt: float,
                steps_per_segment: int, order: int, precision: int, e0: int,
                checkpoint_every: int):
        if g_segments.dim() != 1:
            raise ValueError("g_segments must be 1-D")
        g = g_segments.detach().to("cpu", torch.float64).numpy()
        g_steps = np.repeat(g, steps_per_segment)
        pl, dh, _ = engine.adjoint_forward(n, start_bits, g_steps, dt, order=order,
                                           precision=precision, e0=e0,
                                           checkpoint_every=checkpoint_every)
        engine._adj_token = getattr(engine, "_adj_token", 0) + 1
        ctx.engine = engine
        ctx.token = engine._adj_token
        ctx.nseg = int(g.shape[0])
        ctx.sps = int(steps_per_segment)
        return torch.tensor([pl, dh], dtype=g_segments.dtype, device=g_segments.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        eng = ctx.engine
        if getattr(eng, "_adj_token", None) != ctx.token:
            raise RuntimeError("the engine's adjoint tape belongs to a later forward pass; "
                               "run backward before the next forward on the same engine")
        w = grad_out.detach().to("cpu", torch.float64).numpy()
        grad_steps, drift = eng.adjoint_backward(float(w[0]), float(w[1]))
        eng.last_drift = drift
        gseg = grad_steps.reshape(ctx.nseg, ctx.sps).sum(axis=1)
        g = torch.from_numpy(gseg).to(device=grad_out.device, dtype=grad_out.dtype)
        return (g,) + (None,) * 9


def trotter_observables(g_segments: torch.Tensor, engine, n: int, start_bits: int, dt: float,
                        steps_per_segment: int, order: int = _bk.BK_STRANG,
                        precision: int = _bk.BK_C64, e0: int = -1,
                        checkpoint_every: int = 64) -> torch.Tensor:
    """[P_<(T), <d_H>(T)] after len(g_segments) * steps_per_segment Trotter steps
    from |start_bits>, differentiable in g_segments."""
    return TrotterObservables.apply(g_segments, engine, n, start_bits, dt, steps_per_segment,
                                    order, precision, e0, checkpoint_every)


def optimize_schedule(engine, n: int, start_bits: int, g0: np.ndarray, dt: float,
                      steps_per_segment: int, iters: int = 50, lr: float = 0.01,
                      order: int = _bk.BK_STRANG, precision: int = _bk.BK_C64,
                      checkpoint_every: int = 64, e0: int = -1,
                      device: Optional[str] = None) -> dict:
    """Adam on the segment fields with loss -P_<(T).  Returns the loss history
    and the final fields."""
    g = torch.tensor(np.asarray(g0, dtype=np.float64), requires_grad=True,
                     device=device or "cpu")
    opt = torch.optim.Adam([g], lr=lr)
    losses = []
    for _ in range(iters):
        opt.zero_grad()
        obs = trotter_observables(g, engine, n, start_bits, dt, steps_per_segment, order=order,
                                  precision=precision, e0=e0, checkpoint_every=checkpoint_every)
        loss = -obs[0]
        loss.backward()
        opt.step()
        losses.append(float(loss.detach()))
    return {"loss": np.array(losses), "g": g.detach().cpu().numpy(), "lr": lr}
