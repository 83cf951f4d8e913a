"""The batched discrete adjoint at the workload's own sizes and against the
FP64 CPU reference gradient.

Workload test: the first three members of the 32-member batched-adjoint
workload (each side drawn from {64, 96, 128}), viscous NS3D (Re 1000,
Mach 0.1, Pr 0.72) with characteristic WENO5-JS + Roe, each member with its
own dt from CFL 0.5, 30 RK4 steps.  Per member, dJ_m/d(initial velocity) of
the member's final kinetic energy, with a checkpoint every 10 steps (three
segments, all kept by the forward sweep) and every 5 steps (six segments,
three of them recomputed by the reverse sweep), against central finite
differences of the batched forward along two seeded smooth unit directions.

Small-member test: the same scheme on three small uneven members through
6 steps with a checkpoint every step (three segments recomputed), each
member's gradient against torch autograd through oracle/torch_ref.py on the
CPU.
"""
import dataclasses

import pytest
import torch

from hypar_amd import adjoint, decks

pytestmark = pytest.mark.gpu


def _velocity_problem(ws, device="cuda"):
    """Density, velocity, pressure and dt of every member (velocity is the
    control; density and pressure are held)."""
    rhos, vels, ps, dts = [], [], [], []
    for w in ws:
        u0, dt = decks.prepare(w)
        t = torch.as_tensor(u0, device=device).permute(3, 0, 1, 2)
        rho = t[0].contiguous()
        v = (t[1:4] / rho).contiguous()
        p = ((w.gamma - 1.0) * (t[4] - 0.5 * rho * (v * v).sum(0))).contiguous()
        rhos.append(rho)
        vels.append(v)
        ps.append(p)
        dts.append(dt)
    return rhos, vels, ps, dts


def _direction(v0, dims, seed):
    """Unit direction: the normalised initial velocity plus a normalised
    field of random-phase modes (|k| <= 4) on each component."""
    r = torch.stack([torch.tensor(decks.density_perturbation(dims, seed + j), device=v0.device)
                     for j in range(3)])
    d = v0 / v0.norm() + r / r.norm()
    return d / d.norm()


def test_batched_adjoint_workload_matches_finite_differences():
    from hypar_amd.ensemble import Ensemble
    ws = decks.ensemble_config4(n_members=3)
    dims = [w.dims for w in ws]
    nsteps = 30
    print(f"[gpu] batched adjoint workload: members {dims}, {nsteps} steps", flush=True)
    rhos, vels, ps, dts = _velocity_problem(ws)
    e = Ensemble(dims, "characteristic", "roe", Re=1000.0, Minf=0.1, Pr=0.72)
    e.set_dt(dts)
    dVs = [adjoint.cell_volumes(d) for d in dims]

    def state(vs):
        return torch.cat([adjoint.conserved_state(r, (v[0], v[1], v[2]), p) for r, v, p in zip(rhos, vs, ps)])

    def J(u):
        return [adjoint.kinetic_energy(e.member(u, m), dims[m], dVs[m]) for m in range(len(dims))]

    grads = {}
    for every in (10, 5):
        ctrl = [v.clone().requires_grad_(True) for v in vels]
        uN = adjoint.ensemble_trajectory(state(ctrl), e, nsteps, every)
        grads[every] = torch.autograd.grad(sum(J(uN)), ctrl)
        print(f"[gpu]   checkpoint every {every}: gradients finite "
              f"{all(bool(torch.isfinite(g).all()) for g in grads[every])}", flush=True)
    for m in range(len(dims)):
        diff = float((grads[10][m] - grads[5][m]).norm() / grads[10][m].norm())
        print(f"[gpu]   member {m} {dims[m]}: every 10 vs every 5 rel diff {diff:.2e}", flush=True)
        assert diff <= 1e-13

    h = 1e-4
    for k in range(2):
        ds = [_direction(vels[m], dims[m], 4100 + 10 * m + 3 * k) for m in range(len(dims))]
        with torch.no_grad():
            up = state([v + h * d for v, d in zip(vels, ds)])
            um = state([v - h * d for v, d in zip(vels, ds)])
            e.rk4(up, nsteps)
            e.rk4(um, nsteps)
            jp, jm = J(up), J(um)
        for m in range(len(dims)):
            fd = float((jp[m] - jm[m]) / (2 * h))
            for every in (10, 5):
                ad = float((grads[every][m] * ds[m]).sum())
                rel = abs(fd - ad) / abs(ad)
                print(f"[gpu]   member {m} direction {k} every {every}: adjoint {ad:.12e} fd {fd:.12e} "
                      f"rel {rel:.2e}", flush=True)
                assert rel <= 1e-6
    e.close()


def test_batched_velocity_gradient_matches_torch_autograd():
    from hypar_amd.ensemble import Ensemble
    from oracle import torch_ref as T
    dims = [(12, 10, 14), (10, 12, 8), (14, 8, 10)]
    ws = [dataclasses.replace(decks.ensemble_config4(n_members=1)[0], dims=d, seed=20264 * 1000 + 50 + m, dt=None)
          for m, d in enumerate(dims)]
    nsteps, every = 6, 1
    print(f"[gpu] batched velocity gradient vs torch autograd: members {dims}, {nsteps} steps", flush=True)
    rhos, vels, ps, dts = _velocity_problem(ws, device="cpu")
    e = Ensemble(dims, "characteristic", "roe", Re=1000.0, Minf=0.1, Pr=0.72)
    e.set_dt(dts)
    ctrl = [v.to("cuda").requires_grad_(True) for v in vels]
    u0 = torch.cat([adjoint.conserved_state(r.cuda(), (v[0], v[1], v[2]), p.cuda())
                    for r, v, p in zip(rhos, ctrl, ps)])
    uN = adjoint.ensemble_trajectory(u0, e, nsteps, every)
    got = torch.autograd.grad(sum(adjoint.kinetic_energy(e.member(uN, m), dims[m]) for m in range(len(dims))), ctrl)
    for m, w in enumerate(ws):
        dx = [torch.tensor(decks.dxinv_1d(decks.grid_1d(n))[3:-3]) for n in w.dims]
        sch = dict(interp="characteristic", upwind="roe", Re=w.Re / w.Minf, Pr=w.Pr, weights="js")
        x = vels[m].clone().requires_grad_(True)
        y = adjoint.conserved_state(rhos[m], (x[0], x[1], x[2]), ps[m]).view(5, *reversed(w.dims))
        for _ in range(nsteps):
            y = T.rk4_step(y, dts[m], dx, **sch)
        (ref,) = torch.autograd.grad(T.kinetic_energy(y, dx), x)
        err = float((got[m].cpu() - ref).norm() / ref.norm())
        print(f"[gpu]   member {m} {w.dims}: rel err vs torch autograd {err:.2e}", flush=True)
        assert err <= 1e-11
    e.close()
