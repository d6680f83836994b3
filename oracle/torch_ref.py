"""FP64 reference gradient for the differentiable trajectory.

TEST INFRASTRUCTURE ONLY (see oracle/ljref.py).  A brute-force restatement of
the reference's LJ spline force (physics/lennard_jones_force.cpp:76-128, every
pair and all 27 lattice images as engine/particle_pair_filter.cpp:67-123) and
velocity Verlet (engine/integrator.cpp:62-88), generalised to V = eps v(r /
sigma) with rc = cutoff * sigma, written with torch float64 tensors on the
CPU so that torch's own reverse mode supplies an independent gradient of

    L = (1/n) sum_{t=1..n} [PE_t + rho (T_t + W_t / (3N))]

with forces F(r0) before the first half kick.  Small N only (O(27 N^2)).
"""

import itertools

import torch

IMAGES = torch.tensor(list(itertools.product((-1.0, 0.0, 1.0), repeat=3)), dtype=torch.float64)


def pair_terms(r, L, eps, sigma, cutoff):
    """Forces (N, 3), PE, W of positions r (N, 3) in a cubic box L."""
    n = r.shape[0]
    rc2_red = cutoff * cutoff
    rcm6 = 1.0 / (rc2_red ** 3)
    alpha = -4.0 * rcm6 * (rcm6 - 1.0)
    beta = 12.0 * rcm6 * (2.0 * rcm6 - 1.0)
    sep = (r[:, None, None, :] - r[None, :, None, :]) - IMAGES[None, None, :, :] * L   # (n,n,27,3)
    s = (sep * sep).sum(-1)
    self_mask = torch.zeros(n, n, 27, dtype=torch.bool)
    self_mask[torch.arange(n), torch.arange(n), 13] = True
    inside = (s < (cutoff * sigma) ** 2) & ~self_mask
    s_safe = torch.where(inside, s, torch.ones_like(s))
    x2 = s_safe / (sigma * sigma)
    s6 = 1.0 / (x2 * x2 * x2)
    u = x2 / rc2_red
    e = eps * (4.0 * s6 * (s6 - 1.0) + alpha + beta * (u - 1.0))
    w = eps * (24.0 * s6 * (2.0 * s6 - 1.0) - 2.0 * beta * u)
    e = torch.where(inside, e, torch.zeros_like(e))
    w = torch.where(inside, w, torch.zeros_like(w))
    g = w / s_safe
    F = (g[..., None] * sep).sum((1, 2))
    return F, 0.5 * e.sum(), 0.5 * w.sum()


def trajectory(eps, sigma, r0, v0, L, nsteps, dt, cutoff):
    """Velocity Verlet from (r0, v0) with the forces F(r0) computed first;
    yields (PE, W, T) after each step (positions are not wrapped: the 27
    images cover every pair of a short trajectory)."""
    n = r0.shape[0]
    F, _, _ = pair_terms(r0, L, eps, sigma, cutoff)
    r, v = r0, v0
    for _ in range(nsteps):
        vh = v + 0.5 * dt * F
        r = r + dt * vh
        F, pe, w = pair_terms(r, L, eps, sigma, cutoff)
        v = vh + 0.5 * dt * F
        ke = 0.5 * (v * v).sum()
        yield pe, w, (2.0 / 3.0) * ke / n


def trajectory_loss(eps, sigma, r0, v0, L, nsteps, dt, cutoff, density):
    n = r0.shape[0]
    total = torch.zeros((), dtype=torch.float64)
    for pe, w, T in trajectory(eps, sigma, r0, v0, L, nsteps, dt, cutoff):
        total = total + pe + density * (T + w / (3.0 * n))
    return total / nsteps
