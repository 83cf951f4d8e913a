"""Discrete adjoint of the WENO5 RK4 step: component-wise and
characteristic-wise, Roe and Rusanov upwinding, inviscid and viscous, with
JS, mapped, WENO-Z and YC weights.

The reference has no adjoint, so the checks are (a) the FP64 CPU reference
gradient: torch.autograd through oracle/torch_ref.py, whose forward is pinned
here to the C oracle (and so to the reference), and (b) central finite
differences of the GPU forward along seeded random directions.
"""
import dataclasses

import numpy as np
import pytest
import torch

from hypar_amd import decks
from oracle import torch_ref as T
from oracle.oracle import Oracle

DIMS = (12, 10, 14)


# ------------------------------------------------------------------ CPU
SCHEMES = [("components", "rusanov", -1.0), ("characteristic", "roe", -1.0), ("characteristic", "rusanov", -1.0),
           ("components", "roe", -1.0), ("characteristic", "roe", 1000.0), ("components", "rusanov", 1000.0)]


def _scheme_case(interp, upwind, Re, dims=DIMS, weno_type="js"):
    w = dataclasses.replace(decks.config1(12, 1), dims=dims, dt=None, interp=interp, upwind=upwind, Re=Re,
                            weno_type=weno_type)
    u0, dt = decks.prepare(w)
    dx = [torch.tensor(decks.dxinv_1d(decks.grid_1d(n))[3:-3]) for n in dims]
    scheme = dict(interp=interp, upwind=upwind, Re=Re / w.Minf if Re > 0 else -1.0, Pr=w.Pr, weights=weno_type)
    return w, u0, dt, dx, scheme


# mapped, WENO-Z (Borges), Yamaleev-Carpenter and the optimal (no_limiting) weights
WEIGHT_SCHEMES = [("components", "rusanov", -1.0, "mapped"), ("characteristic", "roe", 1000.0, "borges"),
                  ("characteristic", "rusanov", -1.0, "yc"), ("components", "roe", 1000.0, "yc"),
                  ("characteristic", "roe", -1.0, "mapped"), ("components", "rusanov", 1000.0, "borges"),
                  ("characteristic", "roe", 1000.0, "optimal"), ("components", "rusanov", -1.0, "optimal")]


@pytest.mark.parametrize("interp,upwind,Re", SCHEMES)
def test_torch_reference_matches_oracle(interp, upwind, Re):
    w, u0, dt, dx, sch = _scheme_case(interp, upwind, Re)
    o = Oracle.for_workload(w)
    r = T.to_aos(T.rhs(T.to_soa(u0), dx, **sch))
    assert decks.rel_l2(r, o.rhs(u0)) < 1e-13
    if Re > 0:
        par = T.to_aos(T.parabolic(T.to_soa(u0), dx, sch["Re"], sch["Pr"]))
        assert decks.rel_l2(par, o.parabolic(u0)) < 1e-14
    u2 = T.to_aos(T.rk4_step(T.rk4_step(T.to_soa(u0), dt, dx, **sch), dt, dx, **sch))
    assert decks.rel_l2(u2, o.rk4(u0, dt, 2)) < 1e-14


@pytest.mark.parametrize("interp,upwind,Re,weno", WEIGHT_SCHEMES)
def test_torch_reference_weight_variants_match_oracle(interp, upwind, Re, weno):
    w, u0, dt, dx, sch = _scheme_case(interp, upwind, Re, weno_type=weno)
    o = Oracle.for_workload(w)
    r = T.to_aos(T.rhs(T.to_soa(u0), dx, **sch))
    assert decks.rel_l2(r, o.rhs(u0)) < 1e-13
    u2 = T.to_aos(T.rk4_step(T.rk4_step(T.to_soa(u0), dt, dx, **sch), dt, dx, **sch))
    assert decks.rel_l2(u2, o.rk4(u0, dt, 2)) < 1e-14


@pytest.mark.parametrize("interp,upwind,Re", [("components", "rusanov", -1.0), ("characteristic", "roe", 1000.0)])
def test_torch_reference_gradient_matches_finite_differences(interp, upwind, Re):
    w, u0, dt, dx, sch = _scheme_case(interp, upwind, Re)
    x = T.to_soa(u0).requires_grad_(True)
    J = T.kinetic_energy(T.rk4_step(x, dt, dx, **sch), dx)
    (g,) = torch.autograd.grad(J, x)
    rng = np.random.default_rng(4)
    for _ in range(3):
        d = torch.tensor(rng.normal(size=x.shape))
        d /= d.norm()
        h = 1e-5
        with torch.no_grad():
            jp = T.kinetic_energy(T.rk4_step(x + h * d, dt, dx, **sch), dx)
            jm = T.kinetic_energy(T.rk4_step(x - h * d, dt, dx, **sch), dx)
        fd = float((jp - jm) / (2 * h))
        ad = float((g * d).sum())
        assert abs(fd - ad) <= 1e-6 * abs(ad)


# ------------------------------------------------------------------ GPU
VJP_SCHEMES = SCHEMES + [("components", "roe", 1000.0), ("characteristic", "rusanov", 1000.0)]


def _gpu_solver(w):
    from hypar_amd.solver import NS3D
    return NS3D.from_workload(w)


def _hm(x):
    """(5, nz, ny, nx) tensor -> hm-layout CUDA tensor."""
    return x.detach().reshape(-1).contiguous().cuda()


def _rel(got, ref):
    return float((got - ref).norm() / ref.norm())


@pytest.mark.gpu
@pytest.mark.parametrize("interp,upwind,Re", VJP_SCHEMES)
def test_rhs_vjp_matches_autograd(interp, upwind, Re):
    w, u0, dt, dx, sch = _scheme_case(interp, upwind, Re)
    x = T.to_soa(u0).requires_grad_(True)
    lam = torch.tensor(np.random.default_rng(1).normal(size=x.shape))
    (ref,) = torch.autograd.grad(T.rhs(x, dx, **sch), x, lam)
    print(f"[gpu] rhs vjp {interp}/{upwind} Re={Re} dims={DIMS}", flush=True)
    s = _gpu_solver(w)
    got = s.rhs_vjp(_hm(x), _hm(lam)).cpu().view_as(ref)
    err = _rel(got, ref)
    print(f"[gpu]   rel err {err:.2e}", flush=True)
    assert err <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("interp,upwind,Re", [("characteristic", "roe", 1000.0), ("characteristic", "rusanov", -1.0)])
def test_rhs_vjp_on_ragged_lines(interp, upwind, Re):
    """Sides of 5 and 7 (stencils wrap more than once around a line), sides
    above one and two wavefronts, 300- and 512-cell lines (one to a block of
    320 and 512 threads), and for the viscous tiles (64 x 8 columns, 16
    planes per segment) partial tiles in x, three tiles in y and three z
    segments."""
    for dims in [(5, 7, 6), (70, 5, 9), (130, 6, 5), (9, 19, 37), (300, 5, 6), (6, 5, 512)]:
        w, u0, dt, dx, sch = _scheme_case(interp, upwind, Re, dims)
        x = T.to_soa(u0).requires_grad_(True)
        lam = torch.tensor(np.random.default_rng(7).normal(size=x.shape))
        (ref,) = torch.autograd.grad(T.rhs(x, dx, **sch), x, lam)
        print(f"[gpu] rhs vjp {interp}/{upwind} Re={Re} dims={dims}", flush=True)
        s = _gpu_solver(w)
        err = _rel(s.rhs_vjp(_hm(x), _hm(lam)).cpu().view_as(ref), ref)
        print(f"[gpu]   rel err {err:.2e}", flush=True)
        assert err <= 1e-12


@pytest.mark.gpu
def test_vjp_rejects_lines_longer_than_a_workgroup():
    """One workgroup holds one grid line of at most 512 cells, component-wise
    and characteristic-wise."""
    from hypar_amd import lib
    from hypar_amd.solver import NS3D
    for dims, kw in [((513, 5, 5), {}), ((5, 513, 5), dict(interp="characteristic", upwind="roe"))]:
        s = NS3D(dims, **kw)
        u = torch.ones(5 * s.ncell, dtype=torch.float64, device="cuda")
        with pytest.raises(lib.HMError) as ex:
            s.rhs_vjp(u, u.clone())
        assert ex.value.code == lib.HM_ERR_UNSUPPORTED
        s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("interp,upwind,Re,weno", WEIGHT_SCHEMES)
def test_weight_variant_vjps_match_autograd(interp, upwind, Re, weno):
    """Mapped, WENO-Z, YC and optimal weights: the RHS VJP and the RK4-step VJP against
    torch autograd through the CPU reference (pinned to the oracle above)."""
    w, u0, dt, dx, sch = _scheme_case(interp, upwind, Re, weno_type=weno)
    x = T.to_soa(u0).requires_grad_(True)
    lam = torch.tensor(np.random.default_rng(5).normal(size=x.shape))
    (ref,) = torch.autograd.grad(T.rhs(x, dx, **sch), x, lam)
    (ref_step,) = torch.autograd.grad(T.rk4_step(x, dt, dx, **sch), x, lam)
    print(f"[gpu] rhs and rk4-step vjp {interp}/{upwind} Re={Re} weights={weno}", flush=True)
    s = _gpu_solver(w)
    err = _rel(s.rhs_vjp(_hm(x), _hm(lam)).cpu().view_as(ref), ref)
    g = _hm(lam)
    s.rk4_vjp(_hm(x), dt, g)
    err_step = _rel(g.cpu().view_as(ref_step), ref_step)
    print(f"[gpu]   rel err {err:.2e}, step {err_step:.2e}", flush=True)
    assert err <= 1e-12 and err_step <= 1e-12
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("interp,upwind,Re", VJP_SCHEMES)
def test_rk4_step_vjp_matches_autograd(interp, upwind, Re):
    w, u0, dt, dx, sch = _scheme_case(interp, upwind, Re)
    x = T.to_soa(u0).requires_grad_(True)
    lam = torch.tensor(np.random.default_rng(2).normal(size=x.shape))
    (ref,) = torch.autograd.grad(T.rk4_step(x, dt, dx, **sch), x, lam)
    print(f"[gpu] rk4 step vjp {interp}/{upwind} Re={Re}", flush=True)
    s = _gpu_solver(w)
    g = _hm(lam)
    s.rk4_vjp(_hm(x), dt, g)
    err = _rel(g.cpu().view_as(ref), ref)
    print(f"[gpu]   rel err {err:.2e}", flush=True)
    assert err <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("interp,upwind,Re,weno", [("components", "rusanov", -1.0, "js"),
                                                   ("characteristic", "roe", 1000.0, "js"),
                                                   ("characteristic", "roe", 1000.0, "mapped")])
def test_checkpointed_trajectory_gradient_matches_autograd(interp, upwind, Re, weno):
    """7 steps with a checkpoint every 2 (segments 2, 2, 2, 1: the last three
    kept by the forward sweep, the first recomputed) against torch autograd
    through the unrolled CPU trajectory."""
    from hypar_amd.adjoint import kinetic_energy, trajectory
    w, u0, dt, dx, sch = _scheme_case(interp, upwind, Re, weno_type=weno)
    x = T.to_soa(u0).requires_grad_(True)
    y = x
    for _ in range(7):
        y = T.rk4_step(y, dt, dx, **sch)
    (ref,) = torch.autograd.grad(T.kinetic_energy(y, dx), x)
    print(f"[gpu] trajectory gradient {interp}/{upwind} Re={Re} weights={weno}", flush=True)
    s = _gpu_solver(w)
    xg = _hm(x).requires_grad_(True)
    yN = trajectory(xg, s, dt, 7, 2)
    assert decks.rel_l2(yN.detach().cpu().numpy(), y.detach().reshape(-1).numpy()) < 1e-12
    (got,) = torch.autograd.grad(kinetic_energy(yN, w.dims), xg)
    err = _rel(got.cpu().view_as(ref), ref)
    print(f"[gpu]   rel err {err:.2e}", flush=True)
    assert err <= 1e-11


@pytest.mark.gpu
def test_adjoint_workload_gradient_matches_finite_differences():
    """The adjoint workload: 128^3 inviscid, 50 RK4 steps, checkpoint every 10;
    dJ/d(drho0) along 8 seeded random unit directions vs central differences.

    The directions are seeded fields of the control's own family (random-
    phase modes with |k| <= 4), normalised to unit L2 norm.  White-noise
    directions project onto the smooth gradient with |J'd| ~ 5e-6, where the
    rounding noise of J (~2e-15 absolute, J ~ 15) over 2h already reaches
    1e-6 relative at h = 1e-4; a smooth direction's |J'd| is ~1e3 larger.
    Rusanov's max() and |v| make J piecewise smooth, so h stays small."""
    from hypar_amd.adjoint import cell_volumes, density_control_problem, objective
    from hypar_amd.solver import NS3D
    w = decks.config3(128, 50)
    base, drho = density_control_problem(w)
    s = NS3D(w.dims, "components", "rusanov")
    u0 = torch.stack([base[0] + drho, base[0] * base[1][0], base[0] * base[1][1], base[0] * base[1][2],
                      base[2]])
    dt = decks.cfl_dt(u0.permute(1, 2, 3, 0).cpu().numpy(), w.dims)
    dV = cell_volumes(w.dims)
    x = drho.clone().requires_grad_(True)
    J = objective(x, base, s, dt, 50, 10, dV)
    (g,) = torch.autograd.grad(J, x)
    h = 1e-4
    for k in range(8):
        d = torch.tensor(decks.density_perturbation(w.dims, w.seed + 2000 + k), device="cuda")
        d /= d.norm()
        with torch.no_grad():
            jp = objective(drho + h * d, base, s, dt, 50, 10, dV)
            jm = objective(drho - h * d, base, s, dt, 50, 10, dV)
        fd = float((jp - jm) / (2 * h))
        ad = float((g * d).sum())
        rel = abs(fd - ad) / abs(ad)
        print(f"[gpu] direction {k}: adjoint {ad:.12e} fd {fd:.12e} rel {rel:.2e}", flush=True)
        assert rel <= 1e-6


@pytest.mark.gpu
def test_viscous_characteristic_gradient_matches_finite_differences():
    """The batched-adjoint physics on one grid: viscous (Re 1000, Mach 0.1)
    characteristic WENO5 + Roe at 24^3, 10 RK4 steps with a checkpoint
    every 4; dJ/d(initial velocity) of the final kinetic energy along 4
    seeded smooth unit directions against central differences of the GPU
    forward.  Each direction is the normalised sum of the unit initial
    velocity field and a unit field of random-phase modes (|k| <= 4 on each
    component): a purely random field can be nearly orthogonal to the
    gradient (|J'd| ~ 1e-5 against ~1e-2 along the velocity), and then the
    ~1e-14 rounding noise of J over 2h exceeds 1e-6 of J'd."""
    from hypar_amd import adjoint
    from hypar_amd.solver import NS3D
    dims = (24, 24, 24)
    w, u0, dt, dx, sch = _scheme_case("characteristic", "roe", 1000.0, dims)
    s = NS3D.from_workload(w)
    t = torch.as_tensor(u0, device="cuda").permute(3, 0, 1, 2)
    rho = t[0].contiguous()
    v0 = (t[1:4] / rho).contiguous()
    p = ((w.gamma - 1.0) * (t[4] - 0.5 * rho * (v0 * v0).sum(0))).contiguous()
    dV = adjoint.cell_volumes(dims)

    def J(v, grad=False):
        u = adjoint.conserved_state(rho, (v[0], v[1], v[2]), p)
        if grad:
            uN = adjoint.trajectory(u, s, dt, 10, 4)
        else:
            uN = u.contiguous().clone()
            s.rk4(uN, dt, 10)
        return adjoint.kinetic_energy(uN, dims, dV)

    x = v0.clone().requires_grad_(True)
    (g,) = torch.autograd.grad(J(x, True), x)
    h = 1e-4
    for k in range(4):
        r = torch.stack([torch.tensor(decks.density_perturbation(dims, 777 + 3 * k + j), device="cuda")
                         for j in range(3)])
        d = v0 / v0.norm() + r / r.norm()
        d /= d.norm()
        with torch.no_grad():
            fd = float((J(v0 + h * d) - J(v0 - h * d)) / (2 * h))
        ad = float((g * d).sum())
        rel = abs(fd - ad) / abs(ad)
        print(f"[gpu] direction {k}: adjoint {ad:.12e} fd {fd:.12e} rel {rel:.2e}", flush=True)
        assert rel <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("nsteps", [7, 9])
def test_reverse_sweep_repeats_and_survives_step_vjp(nsteps):
    """A checkpoint every 2 steps and three stage sets.  7 steps: four
    segments, so the first reverse sweep recomputes segment 0 into the set
    that held segment 3.  9 steps: five segments, so a second sweep
    recomputes segment 4 on the solver's stream while the second stream
    recomputes segment 3 (the two share the step buffers and must run one
    after the other).  A second sweep over the same forward (a retain_graph
    backward, or a second objective), and one after hm_rk4_vjp reused stage
    set 0, must recompute what their sets no longer hold: all three give the
    same gradient, equal to torch autograd through the CPU trajectory."""
    from hypar_amd.adjoint import trajectory
    w, u0, dt, dx, sch = _scheme_case("characteristic", "roe", 1000.0)
    x = T.to_soa(u0).requires_grad_(True)
    y = x
    for _ in range(nsteps):
        y = T.rk4_step(y, dt, dx, **sch)
    lam = torch.tensor(np.random.default_rng(11).normal(size=x.shape))
    (ref,) = torch.autograd.grad(y, x, lam)
    print(f"[gpu] repeated reverse sweeps, char/roe viscous, {nsteps} steps, checkpoint every 2", flush=True)
    s = _gpu_solver(w)
    u = _hm(x)
    s.adjoint_forward(u, dt, nsteps, 2)
    assert s.tape_info()["sets"] == 3
    g1 = s.adjoint_reverse(_hm(lam))
    g2 = s.adjoint_reverse(_hm(lam))
    s.rk4_vjp(_hm(x), dt, _hm(lam))
    g3 = s.adjoint_reverse(_hm(lam))
    err = _rel(g1.cpu().view_as(ref), ref)
    print(f"[gpu]   rel err {err:.2e}; sweeps equal: {torch.equal(g1, g2)}, {torch.equal(g1, g3)}", flush=True)
    assert err <= 1e-11
    assert torch.equal(g1, g2) and torch.equal(g1, g3)
    # the same through autograd: two backward passes over one graph
    xg = _hm(x).requires_grad_(True)
    yN = trajectory(xg, s, dt, nsteps, 2)
    (a1,) = torch.autograd.grad(yN, xg, _hm(lam), retain_graph=True)
    (a2,) = torch.autograd.grad(yN, xg, _hm(lam))
    assert torch.equal(a1, g1) and torch.equal(a2, g1)
    s.close()


# (sets, step level) in the order the forward sweep prefers them, and the
# states each holds for 13 steps with a checkpoint every 4 (4 checkpoints):
# checkpoints + sets x (4K or K) + 4 scratch (+3 stage states at step level)
TAPE_PLANS = [((3, False), 56), ((2, False), 40), ((1, False), 24), ((3, True), 23), ((2, True), 19),
              ((1, True), 15)]


@pytest.mark.gpu
def test_tape_schedules_under_budgets_give_the_same_gradient():
    """The forward sweep takes the cheapest tape schedule that fits the
    budget.  13 steps with a checkpoint every 4 (segments 4, 4, 4, 1): with
    the budget set to the bytes of each schedule in turn, that schedule is
    chosen, and the gradient is bit for bit the one of the default tape and
    matches torch autograd through the CPU trajectory (viscous
    characteristic/Roe).  Below the smallest schedule the sweep reports
    HM_ERR_NOMEM with the bytes it needs."""
    from hypar_amd import lib
    w, u0, dt, dx, sch = _scheme_case("characteristic", "roe", 1000.0)
    x = T.to_soa(u0).requires_grad_(True)
    y = x
    for _ in range(13):
        y = T.rk4_step(y, dt, dx, **sch)
    lam = torch.tensor(np.random.default_rng(13).normal(size=x.shape))
    (ref,) = torch.autograd.grad(y, x, lam)
    print("[gpu] tape schedules, char/roe viscous, 13 steps, checkpoint every 4", flush=True)
    s = _gpu_solver(w)
    state_bytes = 5 * int(np.prod(w.dims)) * 8
    s.adjoint_forward(_hm(x), dt, 13, 4)
    info = s.tape_info()
    assert (info["sets"], info["step_level"]) == (3, False)
    work_bytes = info["bytes"] - 56 * state_bytes
    base = s.adjoint_reverse(_hm(lam))
    err = _rel(base.cpu().view_as(ref), ref)
    print(f"[gpu]   default tape {info}: rel err {err:.2e}", flush=True)
    assert err <= 1e-11
    for (sets, step), nstates in TAPE_PLANS:
        s.set_tape_budget(work_bytes + nstates * state_bytes)
        u = _hm(x)
        s.adjoint_forward(u, dt, 13, 4)
        info = s.tape_info()
        g = s.adjoint_reverse(_hm(lam))
        print(f"[gpu]   budget of {nstates} states -> {info}, equal {torch.equal(g, base)}", flush=True)
        assert (info["sets"], info["step_level"]) == (sets, step)
        assert info["bytes"] == work_bytes + nstates * state_bytes
        assert torch.equal(g, base)
    s.set_tape_budget(work_bytes + 15 * state_bytes - 1)
    with pytest.raises(lib.HMError) as err:
        s.adjoint_forward(_hm(x), dt, 13, 4)
    assert err.value.code == lib.HM_ERR_NOMEM and b"needs" in lib.load().hm_last_error()
    s.close()
