"""Differentiation through the RK4 trajectory with the hand-written discrete
adjoint (torch.autograd.Function over the C-ABI).

The forward sweep advances the state on the device and keeps a checkpoint
every `every` steps in HBM; the backward sweep recomputes each checkpoint
segment once with its RK stage states in HBM and steps the adjoint back
(hm_adjoint_forward / hm_adjoint_reverse).  The objective and the control
map are ordinary torch code, so autograd chains them around the trajectory.
"""
from __future__ import annotations

import numpy as np
import torch

from .solver import NS3D


class RK4Trajectory(torch.autograd.Function):
    """u_N = Phi^N(u_0) for an hm-layout state tensor, differentiable in u_0."""

    @staticmethod
    def forward(ctx, u0: torch.Tensor, solver: NS3D, dt: float, nsteps: int, every: int):
        u = u0.detach().clone().contiguous()
        solver.adjoint_forward(u, dt, nsteps, every)
        ctx.solver = solver
        ctx.tape = solver.tape_id
        return u

    @staticmethod
    def backward(ctx, gN: torch.Tensor):
        if ctx.solver.tape_id != ctx.tape:
            raise RuntimeError("RK4Trajectory: the solver recorded another trajectory after this one; "
                               "one solver holds one tape, so run backward before the next forward")
        lam = gN.detach().clone().contiguous()
        ctx.solver.adjoint_reverse(lam)
        return lam, None, None, None, None


def trajectory(u0: torch.Tensor, solver: NS3D, dt: float, nsteps: int, every: int = 10) -> torch.Tensor:
    return RK4Trajectory.apply(u0, solver, dt, nsteps, every)


class EnsembleTrajectory(torch.autograd.Function):
    """Batched u_N = Phi^N(u_0) of all members of an Ensemble (each with its
    own dt), differentiable in the batched u_0."""

    @staticmethod
    def forward(ctx, u0: torch.Tensor, ens, nsteps: int, every: int):
        u = u0.detach().clone().contiguous()
        ens.adjoint_forward(u, nsteps, every)
        ctx.ens = ens
        ctx.tape = ens.tape_id
        return u

    @staticmethod
    def backward(ctx, gN: torch.Tensor):
        if ctx.ens.tape_id != ctx.tape:
            raise RuntimeError("the ensemble's adjoint tape was overwritten by a later forward sweep")
        lam = gN.detach().clone().contiguous()
        ctx.ens.adjoint_reverse(lam)
        return lam, None, None, None


def ensemble_trajectory(u0: torch.Tensor, ens, nsteps: int, every: int = 10) -> torch.Tensor:
    return EnsembleTrajectory.apply(u0, ens, nsteps, every)


def cell_volumes(dims, device="cuda") -> torch.Tensor:
    """dV = 1/(dxinv_x dxinv_y dxinv_z) per cell (the reference's volume
    element, src/HyParFunctions/VolumeIntegral.c:38-39), shaped (nz, ny, nx)."""
    from . import decks
    dx = [torch.tensor(decks.dxinv_1d(decks.grid_1d(n))[3:-3], dtype=torch.float64, device=device)
          for n in dims]
    return 1.0 / (dx[2].view(-1, 1, 1) * dx[1].view(1, -1, 1) * dx[0].view(1, 1, -1))


def kinetic_energy(u: torch.Tensor, dims, dV: torch.Tensor = None) -> torch.Tensor:
    """J = sum over cells of rho |v|^2 / 2 dV for an hm-layout state."""
    nx, ny, nz = dims
    q = u.view(5, nz, ny, nx)
    if dV is None:
        dV = cell_volumes(dims, u.device)
    return (0.5 * (q[1] ** 2 + q[2] ** 2 + q[3] ** 2) / q[0] * dV).sum()


def conserved_state(rho, vel, p, gamma=1.4) -> torch.Tensor:
    """hm-layout conserved state from primitive fields shaped (nz, ny, nx)."""
    E = p / (gamma - 1.0) + 0.5 * rho * (vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)
    return torch.stack([rho, rho * vel[0], rho * vel[1], rho * vel[2], E]).reshape(-1).contiguous()


def density_control_problem(w, seed_offset=1000, device="cuda"):
    """The adjoint workload: base primitive fields of workload w plus a seeded
    density perturbation (amplitude 1e-2, |k| <= 4) as the control.
    Returns (base primitives, drho0) as torch tensors on `device`."""
    from . import decks
    rho, u, v, ww, p = decks.primitive_fields(w)
    drho = decks.density_perturbation(w.dims, w.seed + seed_offset)
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)
    return (t(rho), (t(u), t(v), t(ww)), t(p)), t(drho)


def objective(drho, base, solver: NS3D, dt, nsteps, every=10, dV=None):
    """J(drho) = kinetic energy after nsteps RK4 steps from the base state
    with density rho0 + drho (velocity and pressure held)."""
    rho, vel, p = base
    u0 = conserved_state(rho + drho, vel, p)
    uN = trajectory(u0, solver, dt, nsteps, every)
    return kinetic_energy(uN, solver.dims, dV)
