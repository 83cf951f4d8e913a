"""Differentiable FP64 restatement of the NS3D WENO5 right-hand side and RK4
step in torch (CPU), for reference gradients: component-wise or
characteristic-wise WENO5 (JS, mapped, Z/Borges or YC weights), Rusanov or Roe
upwinding, and the viscous term.

TEST INFRASTRUCTURE ONLY: the adjoint kernels are checked against
torch.autograd through this module (and against central finite
differences); the product never imports it.  Its forward results are pinned
to the C oracle (and so to the reference) by tests/test_adjoint.py.

Follows the same reference files as ns3d_oracle.c: flux and Roe average
include/physicalmodels/navierstokes3d.h:98-169, eigenvectors :254-470,
WENO5 weights include/interpolation.h:265-285 (JS), :314-341 (mapped), :426-447
(Z), :479-500 (YC), interpolants Interp1PrimFifthOrderWENO.c:146-158,
characteristic interpolation Interp1PrimFifthOrderWENOChar.c:86-198 with the
weights of WENOFifthOrderCalculateWeights.c:760-896, Rusanov and Roe
NavierStokes3DUpwind.c:40-124,349-420, flux difference
HyperbolicFunction.c:94-101, viscous term NavierStokes3DParabolicFunction.c:50-325
with FirstDerivativeFourthOrder.c:36-129, RK4 TimeRK.c:126-195.  Arrays are
shaped (5, nz, ny, nx); periodic boundaries are torch.roll.

Mathematically equal forms are used where the reference evaluates one
function several times (the Roe average of the two cells next to an interface
is taken once, not once per stencil orientation); the results agree with the
oracle to rounding.
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


def weno5(m3, m2, m1, p1, p2, g3, g2, g1, q1, q2, eps, weights="js"):
    """WENO5 value at the interface j+1/2 from smoothness stencil m3..p2 and
    interpolated values g3..q2; weights "js", "mapped", "borges" (WENO-Z), "yc"
    or "optimal" (no_limiting)."""
    c1312, c14, s6 = 13.0 / 12.0, 0.25, 1.0 / 6.0
    c = (0.1, 0.6, 0.3)
    b1 = c1312 * (m3 - 2 * m2 + m1) ** 2 + c14 * (m3 - 4 * m2 + 3 * m1) ** 2
    b2 = c1312 * (m2 - 2 * m1 + p1) ** 2 + c14 * (m2 - p1) ** 2
    b3 = c1312 * (m1 - 2 * p1 + p2) ** 2 + c14 * (3 * m1 - 4 * p1 + p2) ** 2
    if weights == "optimal":  # no_limiting: the weights stay at c
        a1, a2, a3 = c
    elif weights in ("js", "mapped"):
        a1, a2, a3 = (ck / (bk + eps) ** 2 for ck, bk in zip(c, (b1, b2, b3)))
        if weights == "mapped":
            s = a1 + a2 + a3
            a1, a2, a3 = ((ak / s) * (ck + ck * ck - 3 * ck * (ak / s) + (ak / s) ** 2)
                          / (ck * ck + (ak / s) * (1.0 - 2.0 * ck)) for ck, ak in zip(c, (a1, a2, a3)))
    else:
        tau = torch.abs(b3 - b1) if weights == "borges" else (m3 - 4 * m2 + 6 * m1 - 4 * p1 + p2) ** 2
        a1, a2, a3 = (ck * (1.0 + (tau / (bk + eps)) ** 2) for ck, bk in zip(c, (b1, b2, b3)))
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


def roe_state(uL, uR, gamma=GAMMA):
    """Velocity and sound speed of the Roe average (navierstokes3d.h:144-169)."""
    rL, vL, eL, PL = prims(uL, gamma)
    rR, vR, eR, PR = prims(uR, gamma)
    HL = 0.5 * (vL[0] ** 2 + vL[1] ** 2 + vL[2] ** 2) + (gamma * PL / rL) / (gamma - 1.0)
    HR = 0.5 * (vR[0] ** 2 + vR[1] ** 2 + vR[2] ** 2) + (gamma * PR / rR) / (gamma - 1.0)
    tL, tR = torch.sqrt(rL), torch.sqrt(rR)
    v = [(tL * vL[k] + tR * vR[k]) / (tL + tR) for k in range(3)]
    H = (tL * HL + tR * HR) / (tL + tR)
    a = torch.sqrt((H - 0.5 * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2)) * (gamma - 1.0))
    return v, a


def eigenvectors(v, a, d, gamma=GAMMA):
    """Left and right eigenvectors of the flux Jacobian along d at the state
    (v, a), as (5, 5, ...) tensors (navierstokes3d.h:288-470)."""
    gm1 = gamma - 1.0
    vx, vy, vz = v
    ek = 0.5 * (vx * vx + vy * vy + vz * vz)
    a2 = a * a
    h0 = a2 / gm1 + ek
    z = torch.zeros_like(a)
    o = torch.ones_like(a)
    vn = v[d]
    # acoustic and entropy rows: entropy row 0, u-a row d+1, u+a row 4
    r0 = [(a2 - gm1 * ek) / a2, gm1 * vx / a2, gm1 * vy / a2, gm1 * vz / a2, -gm1 / a2]
    rm = [(gm1 * ek + a * vn) / (2 * a2)] + [-gm1 * v[k] / (2 * a2) for k in range(3)] + [gm1 / (2 * a2)]
    rp = [(gm1 * ek - a * vn) / (2 * a2)] + [-gm1 * v[k] / (2 * a2) for k in range(3)] + [gm1 / (2 * a2)]
    rm[1 + d] = rm[1 + d] - a / (2 * a2)
    rp[1 + d] = rp[1 + d] + a / (2 * a2)
    if d == 0:
        sa, sb = [vy, z, -o, z, z], [-vz, z, z, o, z]
        rows = [r0, rm, sa, sb, rp]
    elif d == 1:
        sa, sb = [-vx, o, z, z, z], [vz, z, z, -o, z]
        rows = [r0, sa, rm, sb, rp]
    else:
        sa, sb = [vx, -o, z, z, z], [-vy, z, o, z, z]
        rows = [r0, sa, sb, rm, rp]
    L = torch.stack([torch.stack(r) for r in rows])
    # columns of R
    c0 = [o, vx, vy, vz, ek]
    cm = [o, vx, vy, vz, h0 - a * vn]
    cp = [o, vx, vy, vz, h0 + a * vn]
    cm[1 + d] = cm[1 + d] - a
    cp[1 + d] = cp[1 + d] + a
    if d == 0:
        ca, cb = [z, z, -o, z, -vy], [z, z, z, o, vz]
        cols = [c0, cm, ca, cb, cp]
    elif d == 1:
        ca, cb = [z, o, z, z, vx], [z, z, z, -o, -vz]
        cols = [c0, ca, cm, cb, cp]
    else:
        ca, cb = [z, -o, z, z, -vx], [z, z, o, z, vy]
        cols = [c0, ca, cb, cm, cp]
    R = torch.stack([torch.stack(c) for c in cols]).transpose(0, 1)
    return L, R


def matvec(M, x):
    """(5, 5, ...) x (5, ...) -> (5, ...)"""
    return (M * x.unsqueeze(0)).sum(1)


def harten(l, delta=1e-6):
    al = torch.abs(l)
    return torch.where(al < delta, (l * l + delta * delta) / (2 * delta), al)


def interface_flux(u, d, eps=1e-6, gamma=GAMMA, interp="components", upwind="rusanov", weights="js"):
    """WENO5 interface flux at the interface left of every cell along
    direction d (component-wise or characteristic-wise; Rusanov or Roe)."""
    ax = AXIS[d]
    U = [torch.roll(u, shifts=3 - s, dims=ax) for s in range(6)]   # cells i-3 .. i+2
    F = [flux(Us, d, gamma) for Us in U]
    C = [modified(Us, gamma) for Us in U]
    v, a = roe_state(U[2], U[3], gamma)
    if interp == "characteristic" or upwind == "roe":
        L, R = eigenvectors(v, a, d, gamma)
    if interp == "characteristic":
        F = [matvec(L, x) for x in F]
        C = [matvec(L, x) for x in C]
        Um = [matvec(L, x) for x in U]
    else:
        Um = U
    fL = weno5(*F[0:5], *F[0:5], eps, weights)
    fR = weno5(F[5], F[4], F[3], F[2], F[1], F[5], F[4], F[3], F[2], F[1], eps, weights)
    uL = weno5(*Um[0:5], *C[0:5], eps, weights)
    uR = weno5(Um[5], Um[4], Um[3], Um[2], Um[1], C[5], C[4], C[3], C[2], C[1], eps, weights)
    if interp == "characteristic":
        fL, fR, uL, uR = (matvec(R, x) for x in (fL, fR, uL, uR))
    du = 0.5 * (uR - uL)
    if upwind == "roe":
        vn = v[d]
        lam = [harten(vn)] * 5
        lam[1 + d] = harten(vn - a)
        lam[4] = harten(vn + a)
        lam[0] = harten(vn)
        diss = matvec(R, torch.stack(lam) * matvec(L, du))
        return 0.5 * (fL + fR) - diss
    aL, aR = speed(U[2], d, gamma), speed(U[3], d, gamma)
    aA = a + torch.abs(v[d])
    alpha = torch.maximum(torch.maximum(aL, aR), torch.maximum(aR, aA))
    return 0.5 * (fL + fR) - alpha * du


def _d4(fm2, fm1, fp1, fp2):
    return (fm2 - 8 * fm1 + 8 * fp1 - fp2) / 12.0


def _bshape(d, n):
    shape = [1, 1, 1]
    shape[AXIS[d] - 1] = n
    return shape


def parabolic(u, dxinv, Re, Pr=0.72, gamma=GAMMA):
    """Viscous term (NavierStokes3DParabolicFunction.c:50-325) with the
    reference's ghost-layer evaluation: the viscous flux along d at the two
    ghost layers on each side of a line uses the periodic copies of the
    primitive variables, the biased stencil of FirstDerivativeFourthOrder.c:74-93,
    106-125 in the second layer, the ghost dxinv (the nearest interior value),
    and zero derivatives across d.  Re is the solver's (already divided by
    Minf)."""
    gm1 = gamma - 1.0
    rho, v, e, P = prims(u, gamma)
    Q = [v[0], v[1], v[2], gamma * P / rho]
    # interior gradients with periodic neighbours, G[c][e]
    G = []
    for c in range(4):
        row = []
        for ed in range(3):
            ax = AXIS[ed] - 1
            q = Q[c]
            r = _d4(torch.roll(q, 2, ax), torch.roll(q, 1, ax), torch.roll(q, -1, ax), torch.roll(q, -2, ax))
            row.append(r * dxinv[ed].reshape(_bshape(ed, -1)))
        G.append(row)
    par = [torch.zeros_like(rho) for _ in range(5)]
    for d in range(3):
        ax = AXIS[d] - 1
        n = Q[0].shape[ax]
        Qp = [torch.cat([q.narrow(ax, n - 3, 3), q, q.narrow(ax, 0, 3)], dim=ax) for q in Q]

        def at(t, p, ln=1):  # positions p .. p+ln-1 of a padded array (position -3 at index 0)
            return t.narrow(ax, p + 3, ln)

        dd = []
        for q in Qp:
            first = (-3 * at(q, -3) - 10 * at(q, -2) + 18 * at(q, -1) - 6 * at(q, 0) + at(q, 1)) / 12.0
            mid = _d4(at(q, -3, n + 2), at(q, -2, n + 2), at(q, 0, n + 2), at(q, 1, n + 2))
            last = (-at(q, n - 2) + 6 * at(q, n - 1) - 18 * at(q, n) + 10 * at(q, n + 1) + 3 * at(q, n + 2)) / 12.0
            dd.append(torch.cat([first, mid, last], dim=ax))
        dx = dxinv[d]
        dxg = torch.cat([dx[:1], dx[:1], dx, dx[-1:], dx[-1:]]).reshape(_bshape(d, -1))
        Qg = [at(q, -2, n + 4) for q in Qp]
        gr = [[None] * 3 for _ in range(4)]
        for c in range(4):
            for ed in range(3):
                if ed == d:
                    gr[c][ed] = dd[c] * dxg
                else:
                    zpad = torch.zeros_like(G[c][ed].narrow(ax, 0, 2))
                    gr[c][ed] = torch.cat([zpad, G[c][ed], zpad], dim=ax)
        mu = torch.exp(0.76 * torch.log(Qg[3]))
        mr = mu / Re
        kq = mu / Re / gm1 / Pr
        (ux, uy, uz), (vx, vy, vz), (wx, wy, wz), Tg = gr
        if d == 0:
            t = [2.0 / 3.0 * mr * (2 * ux - vy - wz), mr * (uy + vx), mr * (uz + wx)]
        elif d == 1:
            t = [mr * (uy + vx), 2.0 / 3.0 * mr * (-ux + 2 * vy - wz), mr * (vz + wy)]
        else:
            t = [mr * (uz + wx), mr * (vz + wy), 2.0 / 3.0 * mr * (-ux - vy + 2 * wz)]
        fv = t + [Qg[0] * t[0] + Qg[1] * t[1] + Qg[2] * t[2] + kq * Tg[d]]
        di = dx.reshape(_bshape(d, -1))
        for c in range(4):
            f = fv[c]
            par[1 + c] = par[1 + c] + di * _d4(f.narrow(ax, 0, n), f.narrow(ax, 1, n), f.narrow(ax, 3, n),
                                             f.narrow(ax, 4, n))
    return torch.stack(par)


def rhs(u, dxinv, eps=1e-6, gamma=GAMMA, interp="components", upwind="rusanov", Re=-1.0, Pr=0.72,
        weights="js"):
    """-hyperbolic + viscous term; dxinv = (dx, dy, dz) interior inverse
    spacings; Re <= 0 is inviscid (Re already divided by Minf)."""
    hyp = torch.zeros_like(u)
    for d in range(3):
        ax = AXIS[d]
        fI = interface_flux(u, d, eps, gamma, interp, upwind, weights)
        shape = [1, 1, 1, 1]
        shape[ax] = -1
        hyp = hyp + dxinv[d].reshape(shape) * (torch.roll(fI, shifts=-1, dims=ax) - fI)
    if Re > 0:
        return -hyp + parabolic(u, dxinv, Re, Pr, gamma)
    return -hyp


def rk4_step(u, dt, dxinv, eps=1e-6, gamma=GAMMA, **scheme):
    k0 = rhs(u, dxinv, eps, gamma, **scheme)
    k1 = rhs(u + (dt * 0.5) * k0, dxinv, eps, gamma, **scheme)
    k2 = rhs(u + (dt * 0.5) * k1, dxinv, eps, gamma, **scheme)
    k3 = rhs(u + (dt * 1.0) * k2, dxinv, eps, gamma, **scheme)
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
