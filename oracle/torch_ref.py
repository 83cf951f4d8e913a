"""Differentiable FP64 restatement of the inviscid NS3D WENO5 + Rusanov right-hand
side and RK4 step in torch (CPU), for reference gradients.

TEST INFRASTRUCTURE ONLY: the adjoint kernels are checked against
torch.autograd through this module (and against central finite
differences); the product never imports it.  Its forward results are pinned
to the C oracle (and so to the reference) by tests/test_adjoint.py.

Follows the same reference files as ns3d_oracle.c: flux and Roe average
include/physicalmodels/navierstokes3d.h:98-169, WENO5-JS
include/interpolation.h:265-285 and Interp1PrimFifthOrderWENO.c:146-158,
Rusanov NavierStokes3DUpwind.c:349-420, flux difference
HyperbolicFunction.c:94-101, RK4 TimeRK.c:126-195.  Arrays are shaped
(5, nz, ny, nx); periodic boundaries are torch.roll.
"""
from __future__ import annotations

import torch

GAMMA = 1.4
AXIS = {0: 3, 1: 2, 2: 1}  # direction -> tensor dim of a (5, nz, ny, nx) array


def prims(u, gamma=GAMMA):
    rho = u[0]
    vx, vy, vz = u[1] / rho, u[2] / rho, u[3] / rho
    P = (u[4] - 0.5 * rho * (vx * vx + vy * vy + vz * vz)) * (gamma - 1.0)
    return rho, (vx, vy, vz), u[4], P


def flux(u, d, gamma=GAMMA):
    rho, v, e, P = prims(u, gamma)
    vn = v[d]
    f = [rho * vn, rho * vn * v[0], rho * vn * v[1], rho * vn * v[2], (e + P) * vn]
    f[1 + d] = f[1 + d] + P
    return torch.stack(f)


def modified(u, gamma=GAMMA):
    rho, v, e, P = prims(u, gamma)
    e2 = P / (gamma - 1.0) + 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    return torch.cat([u[:4], e2[None]])


def weno5(m3, m2, m1, p1, p2, g3, g2, g1, q1, q2, eps):
    c1312, c14, s6 = 13.0 / 12.0, 0.25, 1.0 / 6.0
    b1 = c1312 * (m3 - 2 * m2 + m1) ** 2 + c14 * (m3 - 4 * m2 + 3 * m1) ** 2
    b2 = c1312 * (m2 - 2 * m1 + p1) ** 2 + c14 * (m2 - p1) ** 2
    b3 = c1312 * (m1 - 2 * p1 + p2) ** 2 + c14 * (3 * m1 - 4 * p1 + p2) ** 2
    a1 = 0.1 / (b1 + eps) ** 2
    a2 = 0.6 / (b2 + eps) ** 2
    a3 = 0.3 / (b3 + eps) ** 2
    s = a1 + a2 + a3
    f1 = (2 * s6) * g3 + (-7 * s6) * g2 + (11 * s6) * g1
    f2 = (-s6) * g2 + (5 * s6) * g1 + (2 * s6) * q1
    f3 = (2 * s6) * g1 + (5 * s6) * q1 + (-s6) * q2
    return (a1 * f1 + a2 * f2 + a3 * f3) / s


def roe_speed(uL, uR, d, gamma=GAMMA):
    """c + |v_d| of the Roe average of two states."""
    rL, vL, eL, PL = prims(uL, gamma)
    rR, vR, eR, PR = prims(uR, gamma)
    HL = 0.5 * (vL[0] ** 2 + vL[1] ** 2 + vL[2] ** 2) + (gamma * PL / rL) / (gamma - 1.0)
    HR = 0.5 * (vR[0] ** 2 + vR[1] ** 2 + vR[2] ** 2) + (gamma * PR / rR) / (gamma - 1.0)
    tL, tR = torch.sqrt(rL), torch.sqrt(rR)
    v = [(tL * vL[k] + tR * vR[k]) / (tL + tR) for k in range(3)]
    H = (tL * HL + tR * HR) / (tL + tR)
    rho = tL * tR
    vsq = v[0] ** 2 + v[1] ** 2 + v[2] ** 2
    P = (H - 0.5 * vsq) * (rho * (gamma - 1.0)) / gamma
    return torch.sqrt(gamma * P / rho) + torch.abs(v[d])


def speed(u, d, gamma=GAMMA):
    rho, v, e, P = prims(u, gamma)
    return torch.sqrt(gamma * P / rho) + torch.abs(v[d])


def interface_flux(u, d, eps=1e-6, gamma=GAMMA):
    """Component-wise WENO5-JS + Rusanov flux at the interface left of every
    cell along direction d."""
    ax = AXIS[d]
    U = [torch.roll(u, shifts=3 - s, dims=ax) for s in range(6)]   # cells i-3 .. i+2
    F = [flux(Us, d, gamma) for Us in U]
    C = [modified(Us, gamma) for Us in U]
    fL = weno5(*F[0:5], *F[0:5], eps)
    fR = weno5(F[5], F[4], F[3], F[2], F[1], F[5], F[4], F[3], F[2], F[1], eps)
    uL = weno5(*U[0:5], *C[0:5], eps)
    uR = weno5(U[5], U[4], U[3], U[2], U[1], C[5], C[4], C[3], C[2], C[1], eps)
    aL, aR = speed(U[2], d, gamma), speed(U[3], d, gamma)
    aA = roe_speed(U[2], U[3], d, gamma)
    alpha = torch.maximum(torch.maximum(aL, aR), torch.maximum(aR, aA))
    return 0.5 * (fL + fR) - alpha * (0.5 * (uR - uL))


def rhs(u, dxinv, eps=1e-6, gamma=GAMMA):
    """-hyperbolic term; dxinv = (dx, dy, dz) interior inverse spacings."""
    hyp = torch.zeros_like(u)
    for d in range(3):
        ax = AXIS[d]
        fI = interface_flux(u, d, eps, gamma)
        shape = [1, 1, 1, 1]
        shape[ax] = -1
        hyp = hyp + dxinv[d].reshape(shape) * (torch.roll(fI, shifts=-1, dims=ax) - fI)
    return -hyp


def rk4_step(u, dt, dxinv, eps=1e-6, gamma=GAMMA):
    k0 = rhs(u, dxinv, eps, gamma)
    k1 = rhs(u + (dt * 0.5) * k0, dxinv, eps, gamma)
    k2 = rhs(u + (dt * 0.5) * k1, dxinv, eps, gamma)
    k3 = rhs(u + (dt * 1.0) * k2, dxinv, eps, gamma)
    return (((u + (dt / 6.0) * k0) + (dt / 3.0) * k1) + (dt / 3.0) * k2) + (dt / 6.0) * k3


def kinetic_energy(u, dxinv):
    """J = sum over cells of rho |v|^2 / 2 dV, dV = 1/(dxinv_x dxinv_y dxinv_z)."""
    dV = 1.0 / (dxinv[2].reshape(-1, 1, 1) * dxinv[1].reshape(1, -1, 1) * dxinv[0].reshape(1, 1, -1))
    ke = 0.5 * (u[1] ** 2 + u[2] ** 2 + u[3] ** 2) / u[0]
    return (ke * dV).sum()


def to_soa(u_aos):
    """(nz, ny, nx, 5) numpy -> (5, nz, ny, nx) float64 tensor."""
    return torch.as_tensor(u_aos).permute(3, 0, 1, 2).contiguous().to(torch.float64)


def to_aos(u_soa):
    return u_soa.detach().permute(1, 2, 3, 0).contiguous().numpy()
