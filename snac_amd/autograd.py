"""torch.autograd binding of the GPU adjoint.

`surface_uplift(theta, problem, param, nsteps)` runs `nsteps` explicit steps
of `problem` on the GPU with the per-element parameter field `theta` and
returns the vertical displacement of the top-surface nodes.  Its backward
is the hand-written adjoint of the step (snh_adjoint_* in the C ABI), so any
misfit built from the uplift in torch differentiates to theta:

    theta = torch.log(refvisc).requires_grad_()
    uplift = surface_uplift(theta, problem, PARAM_LOG_VISCOSITY, 500)
    misfit = 0.5 * ((uplift - target) ** 2).sum()
    misfit.backward()          # theta.grad = d misfit / d ln(viscosity)

theta is ln(refvisc) per element for PARAM_LOG_VISCOSITY (Maxwell) or the
cohesion offset in Pa per element for PARAM_COHESION (plastic).  The
derivative is that of the step with its discrete decisions (dt, damping
branch, plastic branch) frozen at the forward trajectory.
"""
from __future__ import annotations

import copy

import numpy as np
import torch

from . import _abi
from .context import Context
from .problem import Problem


def with_parameter(problem: Problem, param: int, theta: np.ndarray) -> Problem:
    """Copy of problem whose per-element field is set from theta."""
    p = copy.copy(problem)
    p.fields = dict(problem.fields)
    theta = np.asarray(theta, dtype=np.float64).reshape(problem.n_elements)
    if param == _abi.PARAM_LOG_VISCOSITY:
        p.fields[_abi.FIELD_REFVISC] = np.exp(theta)
    elif param == _abi.PARAM_COHESION:
        p.fields[_abi.FIELD_COHESION_OFFSET] = theta.copy()
    else:
        raise ValueError(f"unknown parameter {param}")
    return p


def top_nodes(problem: Problem) -> np.ndarray:
    """Indices of the nodes on the top surface (j = J-1), in node order."""
    I, J, K = problem.nodes
    n = np.arange(problem.n_nodes)
    return n[(n // I) % J == J - 1]


class SurfaceUplift(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, problem, param, nsteps, checkpoint_interval, flags, snapshots):
        p = with_parameter(problem, param, theta.detach().cpu().numpy())
        c = Context(p).Construct().Build().Initialise()
        top = top_nodes(p)
        y0 = c.state(fields=("coord",))["coord"][top, 1]
        ctx.top = top
        ctx.snac = None
        if not ctx.needs_input_grad[0]:
            # no gradient wanted: plain steps, no checkpoints or tape kept
            try:
                c.Execute(nsteps)
                y = c.state(fields=("coord",))["coord"][top, 1]
            finally:
                c.Destroy()
            return torch.as_tensor(y - y0, dtype=theta.dtype, device=theta.device)
        c.adjoint_begin(param, checkpoint_interval, flags, snapshots)
        c.adjoint_forward(nsteps)
        y = c.state(fields=("coord",))["coord"][top, 1]
        ctx.snac = c
        return torch.as_tensor(y - y0, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, grad_uplift):
        c = ctx.snac
        if c is None:
            raise RuntimeError("surface_uplift: the adjoint record is released after the first backward; "
                               "run the forward again for another backward")
        cb = np.zeros((c.N, 3))
        cb[ctx.top, 1] = grad_uplift.detach().cpu().numpy()
        try:
            g = c.adjoint_reverse(cb)
        finally:
            c.adjoint_end()
            c.Destroy()
            ctx.snac = None
        return (torch.as_tensor(g, dtype=grad_uplift.dtype, device=grad_uplift.device),
                None, None, None, None, None, None)


def surface_uplift(theta: torch.Tensor, problem: Problem, param: int, nsteps: int,
                   checkpoint_interval: int = 50, flags: int = 0, snapshots: int = 0) -> torch.Tensor:
    """Uplift of the top-surface nodes after nsteps, differentiable in the
    per-element parameter theta (the adjoint; checkpoint_interval, flags and
    snapshots as for Context.adjoint_begin)."""
    return SurfaceUplift.apply(theta, problem, param, nsteps, checkpoint_interval, flags, snapshots)
