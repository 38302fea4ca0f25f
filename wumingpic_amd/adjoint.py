"""torch.autograd.Function harness over the adjoint C-ABI (include/wpic_adj.h).

`WeibelMagneticEnergy.apply(theta)` with theta = (udrift, uth_e, amp) runs the
2D periodic Weibel beam problem from the seeded state of wpic_ic_beams2d with
Bz += amp sin(2 pi mode x / Lx), returns the final magnetic energy W_B, and its
backward returns dW_B/dtheta from the device adjoint (checkpoint every K steps,
replay in between).  torch carries only the three scalars; all the work runs in
the HIP kernels behind the C-ABI.
"""
from dataclasses import dataclass

import torch

from .pic2d import Pic2D, weibel_charges


@dataclass
class WeibelCase:
    nx: int = 256
    ny: int = 256
    n_ppc: int = 16
    nsteps: int = 200
    every: int = 20
    seed: int = 20261202
    uth_i: float = 0.01
    bz_mode: int = 1
    delt: float = 0.5
    omega_pe: float = 0.1
    mass_ratio: float = 25.0
    cg_tol: float = 1e-13


def make_sim(case, udrift, uth_e, amp):
    q, r = weibel_charges(case.n_ppc, case.omega_pe, case.mass_ratio)
    sim = Pic2D(case.nx, case.ny, q, r, delt=case.delt, capacity=2 * case.n_ppc * case.nx * case.ny)
    sim.set_cg_tol(case.cg_tol)
    sim.ic_beams(case.seed, case.n_ppc, udrift, uth_e, case.uth_i, 0.0)
    if amp != 0.0:
        sim.perturb_bz(amp, case.bz_mode)
    return sim


def magnetic_energy(case, udrift, uth_e, amp):
    """W_B after case.nsteps from the seeded state (forward only)"""
    with make_sim(case, udrift, uth_e, amp) as sim:
        sim.step(case.nsteps)
        return float(sim.energy()[2])


class WeibelMagneticEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, case):
        ud, uth, amp = (float(v) for v in theta.detach().cpu())
        sim = make_sim(case, ud, uth, amp)
        wb = sim.adj_forward(case.nsteps, case.every)
        ctx.sim, ctx.params, ctx.case = sim, (ud, uth), case
        return torch.tensor(wb, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad):
        sim = ctx.sim
        try:
            sim.adj_backward(float(grad))
            g = sim.adj_param_grads(ctx.params[0], ctx.params[1], ctx.case.bz_mode)
        finally:
            sim.close()
        return torch.tensor(g, dtype=torch.float64), None
