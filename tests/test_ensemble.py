"""Batched ensembles (hm_ensemble_*): members of uneven sizes, each with its
own dt, inviscid or viscous, advanced together by persistent launches
replayed from a HIP graph.

Checks: every member against the single-grid path (the same device
arithmetic, so agreement is at rounding level or exact), against the
reference's own ensemble output (golden fixture written by the reference
binary with a shared dt), invariance under the tile shape and grid size the
autotuner chooses between, and graph replays interleaved with eager work.
"""
import dataclasses

import numpy as np
import pytest

from conftest import load_fixture

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-12
DIMS = [(17, 13, 11), (32, 48, 64), (5, 5, 5), (64, 32, 48), (70, 7, 9), (6, 66, 5)]


def _members(dims, interp="components", upwind="rusanov", Re=-1.0, weno="js"):
    from hypar_amd import decks
    ws, u0s, dts = [], [], []
    for m, d in enumerate(dims):
        w = dataclasses.replace(decks.config0(8, 1), dims=tuple(d), interp=interp, upwind=upwind,
                                seed=20262 * 1000 + m, dt=None, Re=Re, Minf=0.1 if Re > 0 else 1.0, weno_type=weno)
        u0, dt = decks.prepare(w)
        ws.append(w)
        u0s.append(u0)
        dts.append(dt)
    return ws, u0s, dts


def _single(w, u0, dt, nsteps):
    from hypar_amd.solver import NS3D
    s = NS3D.from_workload(w)
    u = s.to_device(u0)
    s.rk4(u, dt, nsteps)
    out = s.to_host(u)
    s.close()
    return out


@pytest.mark.parametrize("interp,upwind,Re,weno", [("components", "rusanov", -1.0, "js"),
                                                   ("characteristic", "roe", -1.0, "js"),
                                                   ("characteristic", "roe", 1000.0, "js"),
                                                   ("characteristic", "roe", 1000.0, "mapped"),
                                                   ("components", "rusanov", -1.0, "yc"),
                                                   ("characteristic", "rusanov", -1.0, "borges"),
                                                   ("components", "roe", -1.0, "optimal")])
def test_members_match_single_grid(interp, upwind, Re, weno):
    from hypar_amd import decks
    from hypar_amd.ensemble import Ensemble
    ws, u0s, dts = _members(DIMS, interp, upwind, Re, weno)
    print(f"[gpu] ensemble {interp}/{upwind} Re={Re} weights={weno} dims={DIMS}", flush=True)
    e = Ensemble(DIMS, interp, upwind, Re=Re, Minf=0.1 if Re > 0 else 1.0, weno_type=weno)
    u = e.to_device(u0s)
    e.set_dt(dts)
    e.rk4(u, 3)
    for m, w in enumerate(ws):
        ref = _single(w, u0s[m], dts[m], 3)
        err = decks.rel_l2(e.to_host(u, m), ref)
        print(f"[gpu]   member {m} {w.dims}: rel L2 vs single grid {err:.2e}", flush=True)
        assert err <= 1e-14


def test_reference_ensemble_fixture():
    """The reference's ensemble_simulations run (3 members, shared dt)."""
    from hypar_amd import decks
    from hypar_amd.ensemble import Ensemble
    meta, z = load_fixture("ens3_s4")
    dims = [tuple(s) for s in meta["sizes"]]
    e = Ensemble(dims)
    u = e.to_device([z[f"u0_{i}"] for i in range(len(dims))])
    e.set_dt([meta["dt"]] * len(dims))
    e.rk4(u, meta["steps"])
    for i in range(len(dims)):
        err = decks.rel_l2(e.to_host(u, i), z[f"uN_{i}"])
        print(f"[gpu] reference ensemble member {i} {dims[i]}: rel L2 {err:.2e}", flush=True)
        assert err <= TOL


@pytest.mark.parametrize("interp,upwind,Re,weno", [("components", "rusanov", -1.0, "js"),
                                                   ("characteristic", "roe", -1.0, "js"),
                                                   ("characteristic", "roe", 1000.0, "js"),
                                                   ("characteristic", "roe", 1000.0, "mapped"),
                                                   ("components", "rusanov", -1.0, "borges"),
                                                   ("characteristic", "rusanov", -1.0, "yc"),
                                                   ("components", "roe", -1.0, "optimal")])
def test_one_launch_per_stage_matches_pass_launches(interp, upwind, Re, weno):
    """The stage kernel (shapes 2, 3: the x, y and z items of all members in
    one launch, ordered by dependency counters) gives the same bits as the
    launch-per-pass kernels: each cell's arithmetic is the same, so any
    missed dependency would show as a difference."""
    from hypar_amd.ensemble import Ensemble
    ws, u0s, dts = _members(DIMS, interp, upwind, Re, weno)
    print(f"[gpu] one launch per stage vs per pass: {interp}/{upwind} Re={Re} weights={weno}", flush=True)
    e = Ensemble(DIMS, interp, upwind, Re=Re, Minf=0.1 if Re > 0 else 1.0, weno_type=weno)
    u0 = e.to_device(u0s)
    e.set_dt(dts)
    out = {}
    for shape in (0, 2, 1, 3):
        e.set_tiles(shape)
        u = u0.clone()
        e.rk4(u, 3)
        out[shape] = u
    assert torch.equal(out[2], out[0])
    assert torch.equal(out[3], out[1])
    assert torch.equal(out[0], out[1])
    e.close()


def test_stage_kernel_wait_limit_reports_timeout():
    """A dependency wait that gives up sets a flag the host reports as
    HM_ERR_TIMEOUT, and the kernel still exits.  Forced from the host with a
    wait limit of zero polls: on a single 64^3 member every block holds an
    x item at first, so the first y items find the x count short.  Then the
    normal limit gives the launch-per-pass result again."""
    from hypar_amd import lib
    from hypar_amd.ensemble import Ensemble
    dims = [(64, 64, 64)]
    ws, u0s, dts = _members(dims)
    e = Ensemble(dims)
    u0 = e.to_device(u0s)
    e.set_dt(dts)
    e.set_tiles(0)
    ref = u0.clone()
    e.rk4(ref, 2)
    e.set_tiles(2)
    e.set_wait_limit(0)
    u = u0.clone()
    with pytest.raises(lib.HMError) as err:
        e.rk4(u, 2)
    assert err.value.code == lib.HM_ERR_TIMEOUT
    print(f"[gpu] wait limit 0: {err.value}", flush=True)
    e.set_wait_limit(1 << 21)
    u = u0.clone()
    e.rk4(u, 2)
    assert torch.equal(u, ref)
    e.close()


def test_tile_shapes_and_autotuner_agree():
    from hypar_amd.ensemble import Ensemble
    ws, u0s, dts = _members(DIMS)
    e = Ensemble(DIMS)
    u0 = e.to_device(u0s)
    e.set_dt(dts)
    results = {}
    # shapes 0, 1: a launch per pass; 2, 3: one launch per stage, with the
    # co-resident grid, a grid far smaller (blocks take many items and wait
    # on other members), and one larger than the work list
    cases = [(sh, g) for sh in (0, 1) for g in ((0, 0, 0), (256, 256, 256), (1024, 512, 2048))]
    cases += [(sh, g) for sh in (2, 3) for g in ((0, 0, 0), (7, 0, 0), (100000, 0, 0))]
    for shape, grid in cases:
        e.set_tiles(shape, grid)
        u = u0.clone()
        e.rk4(u, 2)
        results[(shape, grid)] = u
    first = next(iter(results.values()))
    for key, u in results.items():
        assert torch.equal(u, first), key
    keep = u0.clone()
    shape, grid = e.tune(u0, reps=1)
    print(f"[gpu] autotuner chose shape {shape} grid {grid}", flush=True)
    assert torch.equal(u0, keep)  # the tuner works on a copy
    assert (shape, grid) == e.tiles()
    u = u0.clone()
    e.rk4(u, 2)
    assert torch.equal(u, first)


@pytest.mark.parametrize("shape", [0, 2])
def test_graph_replays_between_eager_work(shape):
    """Replay the captured step several times with eager kernels and host
    synchronisation in between; every replay matches eager single-grid steps
    (a launch per pass, and one launch per stage)."""
    from hypar_amd import decks
    from hypar_amd.ensemble import Ensemble
    from hypar_amd.solver import NS3D
    dims = [(16, 12, 20), (24, 16, 8)]
    ws, u0s, dts = _members(dims)
    e = Ensemble(dims)
    u = e.to_device(u0s)
    e.set_dt(dts)
    e.set_tiles(shape)
    singles = [NS3D.from_workload(w) for w in ws]
    refs = [s.to_device(a) for s, a in zip(singles, u0s)]
    for rep in range(4):
        e.rk4(u, 1)
        for s, r, dt in zip(singles, refs, dts):
            s.rk4(r, dt, 1)
        junk = torch.randn(1 << 20, device="cuda", dtype=torch.float64).sum()  # eager work
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        for m, (s, r) in enumerate(zip(singles, refs)):
            err = decks.rel_l2(e.to_host(u, m), s.to_host(r))
            assert err <= 1e-14, (rep, m, err)


def test_viscous_reference_fixtures():
    """Viscous characteristic/Roe members, each with its own dt, against the
    reference's output for the same decks: member 0 after one step, member 1
    (a ragged grid) after three."""
    from hypar_amd import decks
    from hypar_amd.ensemble import Ensemble
    fx = [load_fixture("c1_16_s1"), load_fixture("c1_ragged_s3")]
    dims = [tuple(meta["dims"]) for meta, _ in fx]
    m0 = fx[0][0]
    e = Ensemble(dims, "characteristic", "roe", Re=m0["Re"], Minf=m0["Minf"], Pr=m0["Pr"])
    u = e.to_device([z["u0"] for _, z in fx])
    e.set_dt([meta["dt"] for meta, _ in fx])
    e.rk4(u, 1)
    err = decks.rel_l2(e.to_host(u, 0), fx[0][1]["uN"])
    print(f"[gpu] viscous ensemble member 0 {dims[0]} after 1 step: rel L2 vs reference {err:.2e}", flush=True)
    assert err <= TOL
    e.rk4(u, 2)
    err = decks.rel_l2(e.to_host(u, 1), fx[1][1]["uN"])
    print(f"[gpu] viscous ensemble member 1 {dims[1]} after 3 steps: rel L2 vs reference {err:.2e}", flush=True)
    assert err <= TOL


@pytest.mark.parametrize("interp,upwind,Re,weno", [("components", "rusanov", -1.0, "js"),
                                                   ("characteristic", "roe", 1000.0, "js"),
                                                   ("components", "roe", -1.0, "js"),
                                                   ("characteristic", "roe", 1000.0, "yc"),
                                                   ("components", "rusanov", 1000.0, "js"),
                                                   ("components", "roe", 1000.0, "mapped"),
                                                   ("characteristic", "rusanov", 1000.0, "borges")])
def test_ensemble_adjoint_matches_single_grid(interp, upwind, Re, weno):
    """Batched trajectory gradient (5 steps, checkpoint every step: the last
    three segments kept by the forward sweep, the first two recomputed; each
    member with its own dt) against each member's single-grid adjoint, which
    test_adjoint.py checks against torch autograd.  The last member spans two
    x tiles and two z segments of the viscous kernels."""
    from hypar_amd import decks
    from hypar_amd.adjoint import ensemble_trajectory, kinetic_energy, trajectory
    from hypar_amd.ensemble import Ensemble
    from hypar_amd.solver import NS3D
    dims = [(12, 10, 14), (16, 12, 10), (10, 14, 12), (70, 9, 18)]
    ws, u0s, dts = _members(dims, interp, upwind, Re, weno)
    print(f"[gpu] ensemble adjoint {interp}/{upwind} Re={Re} weights={weno} dims={dims}", flush=True)
    e = Ensemble(dims, interp, upwind, Re=Re, Minf=0.1 if Re > 0 else 1.0, weno_type=weno)
    e.set_dt(dts)
    u0 = e.to_device(u0s).requires_grad_(True)
    uN = ensemble_trajectory(u0, e, 5, 1)
    J = sum(kinetic_energy(e.member(uN, m), dims[m]) for m in range(len(dims)))
    (g,) = torch.autograd.grad(J, u0)
    for m, w in enumerate(ws):
        s = NS3D.from_workload(w)
        x = s.to_device(u0s[m]).requires_grad_(True)
        y = trajectory(x, s, dts[m], 5, 1)
        assert decks.rel_l2(e.member(uN, m).detach().cpu().numpy(), y.detach().cpu().numpy()) <= 1e-14
        (gs,) = torch.autograd.grad(kinetic_energy(y, w.dims), x)
        err = float((e.member(g, m) - gs).norm() / gs.norm())
        print(f"[gpu]   member {m} {w.dims}: gradient rel err vs single grid {err:.2e}", flush=True)
        assert err <= 1e-12
        s.close()


def test_ensemble_adjoint_long_lines_match_autograd():
    """Characteristic/Roe viscous members with 300- and 512-cell lines beside a
    short one: the batched line VJPs take 512-thread blocks, each holding one
    long line or up to 102 five-cell lines.  Two steps, checkpoint every step;
    each member's gradient against torch autograd through the CPU trajectory."""
    from hypar_amd import decks
    from hypar_amd.adjoint import ensemble_trajectory, kinetic_energy
    from hypar_amd.ensemble import Ensemble
    from oracle import torch_ref as T
    dims = [(300, 5, 6), (6, 5, 512), (5, 7, 6)]
    ws, u0s, dts = _members(dims, "characteristic", "roe", 1000.0)
    print(f"[gpu] ensemble adjoint, long lines: char/roe Re=1000 dims={dims}", flush=True)
    e = Ensemble(dims, "characteristic", "roe", Re=1000.0, Minf=0.1)
    e.set_dt(dts)
    u0 = e.to_device(u0s).requires_grad_(True)
    uN = ensemble_trajectory(u0, e, 2, 1)
    (g,) = torch.autograd.grad(sum(kinetic_energy(e.member(uN, m), dims[m]) for m in range(len(dims))), u0)
    for m, w in enumerate(ws):
        dx = [torch.tensor(decks.dxinv_1d(decks.grid_1d(n))[3:-3]) for n in w.dims]
        sch = dict(interp="characteristic", upwind="roe", Re=w.Re / w.Minf, Pr=w.Pr, weights="js")
        x = T.to_soa(u0s[m]).requires_grad_(True)
        y = x
        for _ in range(2):
            y = T.rk4_step(y, dts[m], dx, **sch)
        (ref,) = torch.autograd.grad(T.kinetic_energy(y, dx), x)
        err = float((e.member(g, m).cpu().view_as(ref) - ref).norm() / ref.norm())
        print(f"[gpu]   member {m} {w.dims}: rel err vs torch autograd {err:.2e}", flush=True)
        assert err <= 1e-11
    e.close()


def test_rejections():
    from hypar_amd import lib
    from hypar_amd.ensemble import Ensemble
    with pytest.raises(lib.HMError):
        Ensemble([(4, 8, 8)])

    e = Ensemble([(8, 8, 8)])
    u = torch.zeros(5 * 512, dtype=torch.float64, device="cuda")
    with pytest.raises(lib.HMError) as ex:
        e.rk4(u, 1)  # dt not set
    assert ex.value.code == lib.HM_ERR_ARG
    with pytest.raises(ValueError):
        e.rk4(torch.zeros(10, dtype=torch.float64, device="cuda"), 1)


def test_ensemble_tape_keeps_its_dt_and_repeats():
    """The batched reverse sweep differentiates the recorded trajectory: a
    dt set after the forward sweep does not enter it, and a second sweep
    (five segments for three stage sets, so both recompute) gives the same
    gradient as the first."""
    from hypar_amd.ensemble import Ensemble
    dims = [(12, 10, 14), (16, 12, 10)]
    ws, u0s, dts = _members(dims)
    print(f"[gpu] ensemble tape dt and repeated sweeps dims={dims}", flush=True)
    e = Ensemble(dims)
    e.set_dt(dts)
    u = e.to_device(u0s)
    e.adjoint_forward(u, 5, 1)
    lam = torch.tensor(np.random.default_rng(3).normal(size=u.shape), device="cuda")
    g1 = e.adjoint_reverse(lam.clone())
    e.set_dt([2.0 * d for d in dts])
    g2 = e.adjoint_reverse(lam.clone())
    e.set_dt(dts)
    g3 = e.adjoint_reverse(lam.clone())
    print(f"[gpu]   equal: {torch.equal(g1, g2)}, {torch.equal(g1, g3)}", flush=True)
    assert torch.equal(g1, g2) and torch.equal(g1, g3)
    e.close()


@pytest.mark.gpu
def test_batched_tape_schedules_give_the_same_gradient():
    """The batched adjoint under tape budgets of each schedule (three small
    viscous characteristic/Roe members, 13 steps, a checkpoint every 4): the
    schedule named by the budget is chosen and every gradient equals the
    default tape's bit for bit."""
    from hypar_amd.ensemble import Ensemble
    dims = [(12, 10, 14), (10, 12, 8), (14, 8, 10)]
    ws, u0s, dts = _members(dims, "characteristic", "roe", 1000.0)
    e = Ensemble(dims, "characteristic", "roe", Re=1000.0, Minf=0.1)
    u0 = e.to_device(u0s)
    e.set_dt(dts)
    lam = torch.randn(u0.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(7)).cuda()
    state_bytes = 5 * e.ncell * 8
    u = u0.clone()
    e.adjoint_forward(u, 13, 4)
    info = e.tape_info()
    assert (info["sets"], info["step_level"]) == (3, False)
    work_bytes = info["bytes"] - 56 * state_bytes
    base = e.adjoint_reverse(lam.clone())
    plans = [((3, False), 56), ((2, False), 40), ((1, False), 24), ((3, True), 23), ((2, True), 19),
             ((1, True), 15)]
    for (sets, step), nstates in plans:
        e.set_tape_budget(work_bytes + nstates * state_bytes)
        u = u0.clone()
        e.adjoint_forward(u, 13, 4)
        info = e.tape_info()
        g = e.adjoint_reverse(lam.clone())
        print(f"[gpu] batched tape, budget of {nstates} states -> {info}, equal {torch.equal(g, base)}", flush=True)
        assert (info["sets"], info["step_level"]) == (sets, step)
        assert torch.equal(g, base)
    # the one-launch-per-stage kernel for the taped and the recomputed steps
    # (one set: three of the four segments are recomputed)
    e.set_tiles(2)
    e.set_tape_budget(work_bytes + 24 * state_bytes)
    u = u0.clone()
    e.adjoint_forward(u, 13, 4)
    g = e.adjoint_reverse(lam.clone())
    print(f"[gpu] batched tape with the stage kernel: {e.tape_info()}, equal {torch.equal(g, base)}", flush=True)
    assert torch.equal(g, base)
    e.close()
