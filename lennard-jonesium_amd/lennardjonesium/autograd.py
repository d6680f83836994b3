"""Gradients of time-averaged NVE observables through the GPU engine.

``trajectory_loss(eps, sigma, r0, v0, box, ...)`` returns
    L = (1/n) sum_{t=1..n} [PE_t + rho (T_t + W_t / (3N))]
for n FP64 velocity-Verlet steps of the spline-shifted LJ system
V = eps v(r / sigma), rc = cutoff * sigma, started from (r0, v0) with forces
F(r0).  Its backward pass is the engine's hand-written adjoint
(lj_engine_adjoint: reverse Verlet with force-Jacobian products, states
checkpointed every ``every`` steps and recomputed per segment), wrapped in a
``torch.autograd.Function`` so it composes with other torch code.  PyTorch is
only the wrapper: the forward and adjoint both run in the engine's kernels.
"""

import numpy as np
import torch

from lennardjonesium.engine import Engine, SystemSpec


class _TrajectoryLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps, sigma, r0, v0, box, nsteps, dt, cutoff, density, every):
        n = r0.shape[0]
        spec = SystemSpec(n, box, cutoff=float(cutoff), time_delta=float(dt),
                          epsilon=float(eps), sigma=float(sigma))
        with Engine([spec], precision="fp64") as e:
            e.set_state(0, r0.detach().cpu().double().numpy(), v0.detach().cpu().double().numpy())
            loss, ge, gs, gr, gv = e.adjoint(int(nsteps), float(density), int(every))
        dev, dt_ = r0.device, r0.dtype
        ctx.save_for_backward(torch.tensor(ge, dtype=torch.float64),
                              torch.tensor(gs, dtype=torch.float64),
                              torch.from_numpy(gr), torch.from_numpy(gv))
        ctx.dev, ctx.dtype = dev, dt_
        return torch.tensor(loss, dtype=torch.float64, device=dev)

    @staticmethod
    def backward(ctx, gout):
        ge, gs, gr, gv = ctx.saved_tensors
        g = gout.detach().cpu().double()
        return ((g * ge).to(ctx.dev), (g * gs).to(ctx.dev), (g * gr).to(ctx.dev, ctx.dtype),
                (g * gv).to(ctx.dev, ctx.dtype), None, None, None, None, None, None)


def trajectory_loss(eps, sigma, r0, v0, box, nsteps=2000, dt=0.005, cutoff=2.5, density=0.8,
                    every=100):
    """Differentiable w.r.t. eps, sigma (0-d tensors), r0 and v0 ((N, 3) tensors)."""
    eps = torch.as_tensor(eps, dtype=torch.float64)
    sigma = torch.as_tensor(sigma, dtype=torch.float64)
    return _TrajectoryLoss.apply(eps, sigma, r0, v0, box, nsteps, dt, cutoff, density, every)
