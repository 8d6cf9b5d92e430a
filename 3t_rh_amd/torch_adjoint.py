"""torch.autograd bridge to the reverse mode of the RK3 time loop.

``rollout(solver, u0, params, dts)`` returns u^N after ``len(dts)`` rk3_ssp
steps (src/main.rs:351-373) of the GPU solver with a frozen dt schedule,
as a differentiable function of the initial state u0 (an (nx, ny, 6)
float64 tensor in the reference's AoS layout) and of the physics
parameters ``params`` = (omega_ei, omega_er, kappa_e, kappa_i, kappa_r).
Forward and backward both run on the GPU through the C ABI
(rh3t_adjoint_forward / rh3t_adjoint_backward: checkpoints every
``checkpoint_every`` steps, recompute in between); torch only carries the
host tensors.
"""
import numpy as np
import torch

from . import PARAMS


class RK3Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, params, solver, dts, checkpoint_every):
        solver.set_physics(solver.physics.with_params(params.detach().cpu().numpy()))
        solver.set_state(u0.detach().cpu().numpy(), 0.0)
        adj = solver.adjoint(checkpoint_every)
        adj.forward(np.asarray(dts, dtype=np.float64))
        u_n, _ = solver.get_state()
        ctx.adj = adj
        return torch.from_numpy(u_n)

    @staticmethod
    def backward(ctx, grad):
        lam0, pg = ctx.adj.backward(grad.detach().cpu().contiguous().numpy())
        ctx.adj.close()
        return torch.from_numpy(lam0), torch.from_numpy(pg), None, None, None


def rollout(solver, u0, params, dts, checkpoint_every=5):
    """Differentiable u^N(u0, params); see the module docstring."""
    if params.shape != (len(PARAMS),):
        raise ValueError(f"params must hold {PARAMS}")
    return RK3Rollout.apply(u0, params, solver, dts, checkpoint_every)
