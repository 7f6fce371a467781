"""torch.autograd wrapper of the hand-written adjoint in libasd.

``sqt_of_theta(theta, S0, cfg)`` returns S(q,tau) as a float64 tensor
[nlags, nq, 2] (re, im) for parameters theta = (J, D, B_z, alpha); its backward
is libasd's discrete adjoint of RK4-LLG + S(q,tau) (asd_sqt_vjp, with
checkpoint-recompute), so any torch loss built on S(q,tau) gets dL/dtheta.
Forward and backward both run on the GPU through the C ABI; nothing is
differentiated by torch itself.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import Context, params


@dataclass
class SqtConfig:
    ctx: Context
    L: int
    nsteps: int
    every: int
    T0: int
    nlags: int
    q: np.ndarray
    dt: float = 0.01
    spin_length: float = 1.0
    checkpoint_every: int = 50


class SqtFunction(torch.autograd.Function):
    @staticmethod
    def forward(fctx, theta, S0, cfg: SqtConfig):
        J, D, Bz, alpha = (float(x) for x in theta.detach().cpu().tolist())
        p = params(J, D, Bz, alpha, cfg.dt, cfg.spin_length)
        s, _ = cfg.ctx.sqt_vjp(p, cfg.L, S0, cfg.nsteps, cfg.every, cfg.T0, cfg.nlags, cfg.q,
                               checkpoint_every=cfg.checkpoint_every)
        fctx.save_for_backward(theta)
        fctx.S0, fctx.cfg, fctx.p = S0, cfg, p
        return torch.from_numpy(np.stack([s.real, s.imag], axis=-1)).to(theta.device)

    @staticmethod
    def backward(fctx, gout):
        (theta,) = fctx.saved_tensors
        cfg = fctx.cfg
        g = gout.detach().to("cpu", torch.float64).contiguous().numpy()
        _, grad = cfg.ctx.sqt_vjp(fctx.p, cfg.L, fctx.S0, cfg.nsteps, cfg.every, cfg.T0, cfg.nlags, cfg.q, g=g,
                                  checkpoint_every=cfg.checkpoint_every)
        return torch.as_tensor(grad, dtype=theta.dtype, device=theta.device), None, None


def sqt_of_theta(theta: torch.Tensor, S0: torch.Tensor, cfg: SqtConfig) -> torch.Tensor:
    return SqtFunction.apply(theta, S0, cfg)
