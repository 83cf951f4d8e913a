"""Benchmark: NS3D WENO5-RK4 FP64 cell-updates/s on one MI355X.

Workload (the production-size configuration the metric is quoted on):
viscous NavierStokes3D on a periodic 256^3 grid, characteristic WENO5-JS
with Roe upwinding, 4th-order viscous terms (Re 1600, Pr 0.72, Mach 0.1),
classical RK4 at CFL 0.5, seeded Taylor-Green + random-phase field.

A step is one RK4 step of the whole grid with the state resident in HBM.
Prints one JSON line.  Only one GPU exists on this system: --gpus above 1
exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "NS3D WENO5-RK4 FP64 cell-updates/s at 256³; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# algorithmic HBM bytes per cell-update of a fused 3N RK4 step (5 FP64
# variables = 40 B per array; stage 1 reads u, writes acc and U2 = 120 B;
# stages 2-3 read U_s, u, acc, write acc, U_{s+1} = 200 B each; stage 4
# reads U4, acc, writes u = 120 B).  Halo re-reads, metrics and every
# intermediate (the rhs accumulator between passes, the viscous fluxes and
# primitive variables) count zero.
STEP_BYTES_PER_CELL_UPDATE = 640
# one launch of a stage kernel processes one stage of every cell: its share
# of the step's algorithmic bytes is a quarter of 640 B per cell
STAGE_BYTES_PER_CELL = STEP_BYTES_PER_CELL_UPDATE // 4
# operand bytes per cell one launch of each kernel reads and writes in this
# implementation (not the algorithmic figure; reported beside it), as
# (inviscid, viscous extra):
#   pass_x     read U_s, write R
#   pass_y     read U_s and R, write R
#   pass_z     read U_s, R, u, acc; write acc, U_{s+1}   (+ read par 32 B, write Q 32 B)
#   visc_par   read Q (4 primitives), write par (4 components)
#   prims      read u, write Q
OPERAND_BYTES = {"pass_x": (80, 0), "pass_y": (120, 0), "pass_z": (240, 64), "visc_par": (64, 0),
                 "prims": (72, 0)}


def cpu_baseline(n=64, steps=2):
    """Time the reference CPU path on a bounded sample of the same workload:
    the same scheme on an n^3 periodic grid for `steps` RK4 steps.  Uses the
    reference binary built from its sources (oracle/_ref/HyPar), single
    thread: its OpenMP build shares the characteristic-weight scratch between
    threads.  Falls back to the CPU oracle (a restatement) when the binary is
    absent."""
    from hypar_amd import decks
    w = decks.config1(n, steps)
    ref = os.path.join(ROOT, "oracle", "_ref", "HyPar")
    if os.path.exists(ref):
        with tempfile.TemporaryDirectory() as d:
            decks.write_deck(d, w, screen_op_iter=1)
            env = dict(os.environ, OMP_NUM_THREADS="1")
            r = subprocess.run([ref], cwd=d, capture_output=True, text=True, env=env)
            m = re.search(r"total wctime:\s*([0-9.eE+-]+)", r.stdout)
            if r.returncode == 0 and m:
                t = float(m.group(1))
                return {"value": n ** 3 * steps / t, "unit": "cell-updates/s", "cores": 1,
                        "kind": "reference",
                        "sample": f"reference HyPar (gcc -O3, 1 thread), same scheme on a periodic "
                                  f"{n}^3 grid, {steps} RK4 steps, {t:.2f} s of time integration"}
    from oracle.oracle import Oracle
    u0, dt = decks.prepare(w)
    o = Oracle.for_workload(w)
    t0 = time.time()
    o.rk4(u0, dt, steps)
    t = time.time() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": n ** 3 * steps / t, "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "sample": f"CPU oracle, same scheme on a periodic {n}^3 grid, {steps} RK4 steps, {t:.2f} s"}


def host_cores():
    """Threads this process may use: OMP_NUM_THREADS when set (the GPU box
    sets it to the box's CPU share), else the CPUs of the affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline_components(n=64, steps=2):
    """The reference on all host cores: its component-wise scheme (the
    inviscid WENO5-JS + Rusanov configuration, whose OpenMP loops share no
    scratch between threads) on a periodic n^3 grid for `steps` RK4 steps,
    OMP_NUM_THREADS = host_cores().  None when the reference binary is
    absent."""
    from hypar_amd import decks
    ref = os.path.join(ROOT, "oracle", "_ref", "HyPar")
    if not os.path.exists(ref):
        return None
    w = decks.config0(n, steps)
    cores = host_cores()
    with tempfile.TemporaryDirectory() as d:
        decks.write_deck(d, w, screen_op_iter=1)
        env = dict(os.environ, OMP_NUM_THREADS=str(cores))
        r = subprocess.run([ref], cwd=d, capture_output=True, text=True, env=env)
        m = re.search(r"total wctime:\s*([0-9.eE+-]+)", r.stdout)
        if r.returncode != 0 or not m:
            return None
        t = float(m.group(1))
    return {"value": n ** 3 * steps / t, "unit": "cell-updates/s", "cores": cores, "kind": "reference",
            "affinity_cpus": len(os.sched_getaffinity(0)),
            "sample": f"reference HyPar (gcc -O3, OpenMP, {cores} threads), inviscid component-wise WENO5-JS + "
                      f"Rusanov RK4 on a periodic {n}^3 grid, {steps} RK4 steps, {t:.2f} s of time integration"}


def adjoint_ratio(n=128, nsteps=50, every=10):
    """Adjoint/forward time ratio on the differentiation workload: inviscid
    NS3D on a periodic n^3 grid, component-wise WENO5-JS + Rusanov, RK4 at
    CFL 0.5; J = kinetic energy after nsteps steps, gradient with respect to
    the initial density perturbation, a checkpoint every `every` steps.
    Ratio = (forward with checkpoint stores + seed of dJ/du_N + reverse
    sweep with segment recomputation) / (plain forward of the same steps),
    both timed with events on the solver's stream."""
    import torch

    from hypar_amd import adjoint, decks
    from hypar_amd.solver import NS3D
    w = decks.config3(n, nsteps)
    base, drho = adjoint.density_control_problem(w)
    rho, vel, p = base
    u0 = adjoint.conserved_state(rho + drho, vel, p)
    dt = decks.cfl_dt(u0.view(5, n, n, n).permute(1, 2, 3, 0).cpu().numpy(), w.dims)
    s = NS3D(w.dims, "components", "rusanov")
    dV = adjoint.cell_volumes(w.dims)
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    u = u0.clone()
    s.rk4(u, dt, 2)  # warm
    u = u0.clone()
    torch.cuda.synchronize()
    ev[0].record(stream)
    s.rk4(u, dt, nsteps)
    ev[1].record(stream)
    v = u0.clone()
    ev[2].record(stream)
    s.adjoint_forward(v, dt, nsteps, every)
    uN = v.detach().requires_grad_(True)
    (lam,) = torch.autograd.grad(adjoint.kinetic_energy(uN, w.dims, dV), uN)
    lam = lam.contiguous()
    s.adjoint_reverse(lam)
    ev[3].record(stream)
    torch.cuda.synchronize()
    fwd, adj = ev[0].elapsed_time(ev[1]), ev[2].elapsed_time(ev[3])
    same = bool(torch.equal(u, v))
    s.close()
    return {"workload": f"inviscid NS3D periodic {n}^3, WENO5-JS components + Rusanov, RK4, {nsteps} steps, "
                        f"checkpoint every {every}, dJ/du0 of the final kinetic energy",
            "forward_ms": fwd, "adjoint_ms": adj, "ratio": adj / fwd,
            "forward_states_agree": same}


def ensemble_throughput(nsteps=10):
    """Batched-ensemble throughput: 64 inviscid NS3D members on periodic
    grids whose sides are drawn from {32, 48, 64, 96, 128}, each with its own
    random-phase state and its own dt (CFL 0.5); component-wise WENO5-JS +
    Rusanov, RK4.  Tile shape and persistent grid are autotuned first; the
    timed region replays the captured RK4 step nsteps times.  Beside it, the
    same members advanced one after another through the single-grid path
    (the reference's serial member loop), timed over the same steps."""
    import torch

    from hypar_amd import decks
    from hypar_amd.ensemble import Ensemble
    from hypar_amd.solver import NS3D
    ws = decks.ensemble_config2()
    u0s, dts = [], []
    for w in ws:
        u0, dt = decks.prepare(w)
        u0s.append(u0)
        dts.append(dt)
    e = Ensemble([w.dims for w in ws])
    u = e.to_device(u0s)
    e.set_dt(dts)
    shape, grid = e.tune(u, reps=2)
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # both launch structures on a scratch copy: one launch per direction pass
    # (shapes 0, 1 at the tuned grid) and one per RK stage (shapes 2, 3)
    per_shape = {}
    for sh in (0, 1, 2, 3):
        e.set_tiles(sh, grid if sh < 2 and sh == shape else (0, 0, 0))
        w = u.clone()
        e.rk4(w, 1)
        torch.cuda.synchronize()
        ev0.record(stream)
        e.rk4(w, 3)
        ev1.record(stream)
        torch.cuda.synchronize()
        per_shape[sh] = ev0.elapsed_time(ev1) / 3
        del w
    e.set_tiles(shape, grid)
    e.rk4(u, 1)  # capture + warm
    torch.cuda.synchronize()
    ev0.record(stream)
    e.rk4(u, nsteps)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1)
    ncell = e.ncell
    finite = bool(torch.isfinite(u).all())
    e.close()
    del u
    torch.cuda.empty_cache()
    serial = 0.0
    for w, u0, dt in zip(ws, u0s, dts):
        s = NS3D.from_workload(w)
        v = s.to_device(u0)
        s.rk4(v, dt, 1)
        torch.cuda.synchronize()
        ev0.record(stream)
        s.rk4(v, dt, nsteps)
        ev1.record(stream)
        torch.cuda.synchronize()
        serial += ev0.elapsed_time(ev1)
        s.close()
        del v
    torch.cuda.empty_cache()
    return {"workload": "64 inviscid NS3D members, sides from {32,48,64,96,128}, own dt each (CFL 0.5), "
                        "WENO5-JS components + Rusanov, RK4",
            "members": len(ws), "cells": ncell, "steps": nsteps, "ms_per_step": ms / nsteps,
            "value": ncell * nsteps / (ms / 1e3), "unit": "cell-updates/s",
            "tiles": {"shape": shape, "grid": list(grid),
                      "launches_per_stage": 1 if shape >= 2 else 3},
            "ms_per_step_by_shape": {"per_pass_4wf": per_shape[0], "per_pass_2wf": per_shape[1],
                                     "per_stage_4wf": per_shape[2], "per_stage_2wf": per_shape[3]},
            "finite": finite,
            "serial_members_ms_per_step": serial / nsteps,
            "speedup_vs_serial_members": serial / ms}


def ensemble_adjoint(nsteps=30, every=10):
    """Batched adjoint: 32 viscous NS3D members (Re 1000, Mach 0.1, Pr 0.72)
    on periodic grids whose sides are drawn from {64, 96, 128}, each with
    its own Taylor-Green-type state and its own dt (CFL 0.5);
    characteristic WENO5-JS + Roe, RK4.  Per-member gradient of the final
    kinetic energy with respect to the member's initial velocity field,
    through nsteps forward + adjoint steps with a checkpoint every `every`
    steps, all members advanced by the batched launches with the tiles the
    autotuner picks (reported under "tiles": the stage kernel is one launch
    per RK stage, after the viscous-term launch).  Ratio = (batched
    forward with checkpoints + reverse sweep + the chain rule onto the
    velocity) / (batched plain forward of the same steps)."""
    import torch

    from hypar_amd import adjoint, decks
    from hypar_amd.ensemble import Ensemble
    ws = decks.ensemble_config4(n_iter=nsteps)
    w0 = ws[0]
    rhos, vels, ps, dts, dVs = [], [], [], [], []
    for w in ws:
        u0, dt = decks.prepare(w)
        t = torch.as_tensor(u0, device="cuda").permute(3, 0, 1, 2)
        rho = t[0].contiguous()
        v = (t[1:4] / rho).contiguous()
        p = (w.gamma - 1.0) * (t[4] - 0.5 * rho * (v * v).sum(0))
        rhos.append(rho)
        vels.append(v)
        ps.append(p.contiguous())
        dts.append(dt)
        dVs.append(adjoint.cell_volumes(w.dims))
    dims = [w.dims for w in ws]
    e = Ensemble(dims, "characteristic", "roe", Re=w0.Re, Minf=w0.Minf, Pr=w0.Pr)
    e.set_dt(dts)

    def state(vs):
        return torch.cat([adjoint.conserved_state(r, (v[0], v[1], v[2]), p) for r, v, p in zip(rhos, vs, ps)])

    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    with torch.no_grad():
        u0 = state(vels).contiguous()
    shape, grid = e.tune(u0, reps=2)  # the forward, taped and recomputed steps all use the tuned tiles
    u = u0.clone()
    e.rk4(u, 1)  # capture + warm
    u = u0.clone()
    torch.cuda.synchronize()
    ev[0].record(stream)
    e.rk4(u, nsteps)
    ev[1].record(stream)
    ctrl = [v.clone().requires_grad_(True) for v in vels]
    torch.cuda.synchronize()
    ev[2].record(stream)
    uN = adjoint.ensemble_trajectory(state(ctrl), e, nsteps, every)
    J = sum(adjoint.kinetic_energy(e.member(uN, m), dims[m], dVs[m]) for m in range(len(dims)))
    grads = torch.autograd.grad(J, ctrl)
    ev[3].record(stream)
    torch.cuda.synchronize()
    fwd, adj = ev[0].elapsed_time(ev[1]), ev[2].elapsed_time(ev[3])
    out = {"workload": f"{len(ws)} viscous NS3D members (Re 1000, Mach 0.1), sides from {{64,96,128}}, own dt each "
                       f"(CFL 0.5), characteristic WENO5-JS + Roe, RK4, {nsteps} steps, checkpoint every {every}, "
                       "per-member dJ/d(initial velocity) of the final kinetic energy",
           "members": len(ws), "cells": e.ncell,
           "tiles": {"shape": shape, "grid": list(grid), "launches_per_stage": (1 if shape >= 2 else 3) + 1},
           "forward_ms": fwd, "adjoint_ms": adj, "ratio": adj / fwd,
           "forward_states_agree": bool(torch.equal(u, uN.detach())),
           "gradients_finite": bool(all(torch.isfinite(g).all() for g in grads))}
    e.close()
    del u, uN, grads, u0
    torch.cuda.empty_cache()
    return out


def dns_ssprk3(n=256, nsteps=10, warmup=2):
    """The scheme of the reference's published GPU example deck
    (DNS_IsotropicTurbulenceDecay_CUDA: viscous NS3D, Re 333.33, Mach 0.3,
    component-wise mapped WENO5 + Roe, SSPRK3, dt 0.005) at n^3 on a seeded
    Taylor-Green-type field, timed with events over nsteps fused SSPRK3
    steps.  The published figure (82.421 s for 1000 steps of 256^3 on 64
    V100s, out.log of that example) is printed beside it as context: a
    different machine and initial field."""
    import torch

    from hypar_amd import decks
    from hypar_amd.solver import NS3D
    w = decks.config_dns(n, nsteps)
    u0, dt = decks.prepare(w)
    s = NS3D.from_workload(w)
    u = s.to_device(u0)
    del u0
    s.ssprk3(u, dt, warmup)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    s.ssprk3(u, dt, nsteps)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / nsteps
    cfl = s.cfl(u, dt)
    s.close()
    del u
    torch.cuda.empty_cache()
    return {"workload": f"viscous NS3D periodic {n}^3 (Re 333.33, Mach 0.3, Pr 0.72), component-wise mapped "
                        f"WENO5 + Roe, SSPRK3, dt {dt}, seeded Taylor-Green-type field",
            "s_per_step": ms / 1e3, "cell_steps_per_s": n ** 3 / (ms / 1e3), "steps": nsteps, "final_cfl": cfl,
            "published_reference_context": {"s_per_step": 82.421 / 1000, "hardware": "64x NVIDIA V100 (64 MPI ranks)",
                                            "source": "Examples/3D/NavierStokes3D/DNS_IsotropicTurbulenceDecay_CUDA/"
                                                      "out.log"}}


def copy_probe(mib=1024, reps=3):
    """A plain device copy of known size (torch copy_ of `mib` MiB), the
    calibration run for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * 1024 * 1024 // 8
    a = torch.rand(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8,
                      "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-ensemble-adjoint", action="store_true")
    ap.add_argument("--no-dns", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py: this system has one GPU; --gpus {args.gpus} is not supported "
              f"(no multi-GPU path is built; every figure is for one GPU).", file=sys.stderr)
        sys.exit(2)

    import torch

    from hypar_amd import decks
    from hypar_amd.solver import NS3D

    torch.cuda.set_device(0)
    w = decks.config1(args.n, args.steps)
    u0, dt = decks.prepare(w)
    s = NS3D.from_workload(w)
    u = s.to_device(u0)
    del u0
    ncell = s.ncell

    s.rk4(u, dt, args.warmup)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    s.rk4(u, dt, args.steps)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    ms_per_step = ms / args.steps
    value = ncell * args.steps / (ms / 1e3)
    cfl = s.cfl(u, dt)  # also a NaN/Inf check of the final state

    # per-kernel durations (HIP events on the solver's stream)
    kt = s.profile_kernels(u, reps=5)
    per_step = {k: 4 * v for k, v in kt.items() if k != "prims"}  # once per stage
    dom = max(per_step, key=per_step.get)
    base, visc_extra = OPERAND_BYTES[dom]
    bytes_launch = STAGE_BYTES_PER_CELL * ncell
    achieved = bytes_launch / (kt[dom] / 1e3) / 1e9
    prof = {}
    pfile = os.path.join(ROOT, "profiles", f"counters_{args.n}.json")
    if os.path.exists(pfile):  # rocprofv3 PMC passes of this workload (see the file)
        prof = json.load(open(pfile))
    traffic = prof.get("hbm_bytes_per_launch", {}).get(dom)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "avg_launch_ms": kt[dom], "algorithmic_bytes_per_launch": bytes_launch,
                "algorithmic_bytes_basis": f"{STEP_BYTES_PER_CELL_UPDATE}/4 B per cell per stage launch",
                "operand_bytes_per_launch": (base + visc_extra) * ncell,
                "operand_GBs": (base + visc_extra) * ncell / (kt[dom] / 1e3) / 1e9}
    flops = prof.get("fp64_flops_per_launch", {}).get(dom)
    roofline_fp64 = None
    if flops:
        tf = flops / (kt[dom] / 1e3) / 1e12
        roofline_fp64 = {"bound": "fp64-valu", "kernel": dom, "achieved": tf, "peak": FP64_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": tf / FP64_PEAK_TFLOPS, "flops_per_launch": flops}
    step_bw = value * STEP_BYTES_PER_CELL_UPDATE / 1e9
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Taylor-Green + random-phase modes |k|<=8)",
        "config": {"workload": f"viscous NS3D periodic {args.n}^3, characteristic WENO5-JS + Roe, "
                               f"4th-order viscous (Re 1600, Pr 0.72, Mach 0.1), RK4 CFL 0.5",
                   "grid": [args.n] * 3, "dt": dt, "final_cfl": cfl},
        "roofline": roofline,
        "roofline_fp64": roofline_fp64,
        "step_roofline": {"bytes_per_cell_update": STEP_BYTES_PER_CELL_UPDATE,
                          "achieved_GBs": step_bw, "peak_GBs": HBM_PEAK_GBS,
                          "frac": step_bw / HBM_PEAK_GBS},
        "kernel_ms_per_step": per_step,
        # the kernels timed one at a time (each launch alone on the stream);
        # inside the step consecutive launches overlap at their edges, so
        # this sum exceeds ms_per_step
        "kernel_ms_sum_isolated": sum(per_step.values()),
    }
    s.close()
    del u
    torch.cuda.empty_cache()
    if not args.no_adjoint:
        line["adjoint"] = adjoint_ratio()
    if not args.no_ensemble:
        line["ensemble"] = ensemble_throughput()
    if not args.no_ensemble_adjoint:
        line["ensemble_adjoint"] = ensemble_adjoint()
    if not args.no_dns:
        line["dns_ssprk3"] = dns_ssprk3()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
        line["cpu_baseline_components_all_cores"] = cpu_baseline_components()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
