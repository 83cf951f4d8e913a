"""Write golden vectors for the NS3D WENO5-RK4 path by running the reference.

Runs the reference HyPar binary built from the reference sources by
``oracle/Makefile.ref`` (into ``oracle/_ref/HyPar``) on decks written by
``hypar_amd.decks`` and stores inputs and outputs as small ``.npz`` fixtures
in this directory.  Run it in the build container (the reference is not
present on GPU machines):

    python tests/golden/make_golden.py            # all fixtures
    python tests/golden/make_golden.py c0_16_s2   # one fixture

The reference runs single-threaded: its OpenMP build shares the eigenvector
scratch arrays of the characteristic WENO weights between threads
(src/InterpolationFunctions/WENOFifthOrderCalculateWeights.c:796-833 declares
``L`` and ``uavg`` outside the ``omp parallel for`` and does not list them as
private), so multi-threaded characteristic runs are not reproducible.

Small grids store the full initial and final interior states.  Large grids
store the workload (seed, sizes, dt) plus checksums and a strided subsample of
the final state; the initial state is regenerated from the seed.

Every deck runs with ``screen_op_iter 1``, and the fixture keeps the CFL and
norm the reference prints after each step (src/TimeIntegration/
TimePrintStep.c:28-33, ``CFL=%1.3E`` and ``norm=%1.4E``) as the printed
strings: the CFL of the state before the step (TimePreStep.c:93-97,
NavierStokes3DComputeCFL.c:16-55; for ensembles the member loop leaves the
last member's value) and the RMS of the step's change (TimePostStep.c:43-63).
"""
from __future__ import annotations

import dataclasses
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from hypar_amd import decks  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "HyPar")


def _workload(kind: str, dims, n_iter: int) -> decks.Workload:
    # "<scheme>+<weights>" selects a WENO weight formulation other than JS
    kind, _, weno = kind.partition("+")
    if weno:
        return dataclasses.replace(_workload(kind, dims, n_iter), weno_type=weno)
    if kind == "c0_ssprk3":
        return dataclasses.replace(decks.config0(dims[0], n_iter), dims=tuple(dims), n_iter=n_iter,
                                   dt=None, time_scheme_type="ssprk3")
    # "<c0|c1>_rk<type>": the other explicit tableaux of solver.inp (1fe, 22, 33)
    for base in ("c0", "c1"):
        if kind.startswith(base + "_rk"):
            return dataclasses.replace(_workload(base, dims, n_iter), time_scheme_type=kind[len(base) + 3:])
    if kind == "dns":
        return dataclasses.replace(decks.config_dns(dims[0], n_iter), dims=tuple(dims), dt=None)
    if kind == "c0":
        w = decks.config0(dims[0], n_iter)
    elif kind == "c1":
        w = decks.config1(dims[0], n_iter)
    else:
        raise ValueError(kind)
    return dataclasses.replace(w, dims=tuple(dims), n_iter=n_iter, dt=None)


# name -> (scheme, dims, steps, store_full)
FIXTURES = {
    "c0_16_s1": ("c0", (16, 16, 16), 1, True),
    "c0_16_s10": ("c0", (16, 16, 16), 10, True),
    "c1_16_s1": ("c1", (16, 16, 16), 1, True),
    "c1_16_s10": ("c1", (16, 16, 16), 10, True),
    "c0_ragged_s3": ("c0", (20, 12, 16), 3, True),
    "c1_ragged_s3": ("c1", (12, 16, 20), 3, True),
    "c0_64_s100": ("c0", (64, 64, 64), 100, False),
    "c1_32_s20": ("c1", (32, 32, 32), 20, False),
    "c1_64_s200": ("c1", (64, 64, 64), 200, False),
    "c0_ssprk3_16_s4": ("c0_ssprk3", (16, 16, 16), 4, True),
    # the scheme of the reference's published GPU deck (viscous, mapped
    # component-wise WENO5 + Roe, SSPRK3), dt from CFL 0.5
    "dns_16_s4": ("dns", (16, 16, 16), 4, True),
    "dns_ragged_s3": ("dns", (12, 20, 16), 3, True),
    # the tableaux without a fused step: HyPar's stage loop around the
    # library's right-hand side
    "c0_rk1fe_16_s2": ("c0_rk1fe", (16, 16, 16), 2, True),
    "c0_rk22_16_s3": ("c0_rk22", (16, 16, 16), 3, True),
    "c1_rk33_16_s3": ("c1_rk33", (16, 16, 16), 3, True),
    # mapped (Henrick et al.), Z (Borges et al.) and Yamaleev-Carpenter weights
    "c0_mapped_16_s4": ("c0+mapped", (16, 16, 16), 4, True),
    "c0_borges_16_s4": ("c0+borges", (16, 16, 16), 4, True),
    "c0_yc_16_s4": ("c0+yc", (16, 16, 16), 4, True),
    "c1_mapped_16_s4": ("c1+mapped", (16, 16, 16), 4, True),
    "c1_borges_16_s4": ("c1+borges", (16, 16, 16), 4, True),
    "c1_yc_16_s4": ("c1+yc", (16, 16, 16), 4, True),
    # no_limiting 1: the weights stay at their optimal values
    "c0_optimal_16_s4": ("c0+optimal", (16, 16, 16), 4, True),
    "c1_optimal_16_s4": ("c1+optimal", (16, 16, 16), 4, True),
}

# fixtures whose decks turn on the conservation check; they also record the
# reference's printed initial volume integral and its conservation.dat
CONSERVATION = {
    "c1_16_s4_cons": ("c1", (16, 16, 16), 4, True),
    "c0_ragged_s3_cons": ("c0", (20, 12, 16), 3, True),
}
FIXTURES.update(CONSERVATION)

ENSEMBLE = {
    # three members of uneven sizes, shared dt and n_iter
    "ens3_s4": ([(16, 12, 20), (12, 20, 16), (20, 16, 12)], 4),
}


def run_reference(deck_dir: str, want_stdout: bool = False):
    env = dict(os.environ, OMP_NUM_THREADS="1")
    t = time.time()
    r = subprocess.run([REF_BIN], cwd=deck_dir, capture_output=True, text=True, env=env)
    if r.returncode != 0:
        raise RuntimeError(f"reference failed in {deck_dir}:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    wall = time.time() - t
    return (wall, r.stdout) if want_stdout else wall


def conservation_outputs(deck_dir: str, stdout: str) -> dict:
    """The initial volume integral the reference prints
    (src/Simulation/InitialSolution.c:127-132) and the conservation errors of
    its conservation.dat (src/Simulation/SimulationWriteErrors.c:83-89)."""
    lines = stdout.splitlines()
    i = next(k for k, l in enumerate(lines) if l.startswith("Volume integral of the initial solution"))
    vi = [float(lines[i + 1 + v].split()[1]) for v in range(5)]
    with open(os.path.join(deck_dir, "conservation.dat")) as f:
        fields = f.read().split()
    return {"volume_integral_initial": vi, "conservation_dat": fields}


def step_trace(stdout: str) -> dict:
    """Per-step printed CFL and norm strings of a run with screen_op_iter 1,
    in both print formats of TimePrintStep.c:23-33 (one line per step, or
    the ensemble block of indented lines)."""
    its = [int(m) for m in re.findall(r"iter=\s*(\d+)", stdout)]
    cfl = re.findall(r"CFL=([0-9.Ee+-]+)", stdout)
    norm = re.findall(r"norm=([0-9.Ee+-]+)", stdout)
    if not (len(its) == len(cfl) == len(norm)) or its != list(range(1, len(its) + 1)):
        raise RuntimeError("could not parse the reference's step lines")
    return {"cfl": cfl, "norm": norm}


def summary(u: np.ndarray) -> dict:
    return {"sum": [float(v) for v in u.reshape(-1, 5).sum(axis=0)],
            "l2": [float(v) for v in np.sqrt((u.reshape(-1, 5) ** 2).sum(axis=0))]}


def subsample(u: np.ndarray, step: int = 4) -> np.ndarray:
    return np.ascontiguousarray(u[::step, ::step, ::step, :])


def make_single(name: str):
    kind, dims, steps, full = FIXTURES[name]
    w = _workload(kind, dims, steps)
    cons = name in CONSERVATION
    extra = {}
    with tempfile.TemporaryDirectory() as d:
        u0 = decks.write_deck(d, w, screen_op_iter=1, conservation_check="yes" if cons else "no")
        wall, out = run_reference(d, want_stdout=True)
        trace = step_trace(out)
        assert len(trace["cfl"]) == steps
        _, _, uN = decks.read_op_bin(os.path.join(d, "op.bin"))
        if cons:
            extra = conservation_outputs(d, out)
    meta = {"name": name, "scheme": kind, "dims": list(dims), "steps": steps, "dt": w.dt,
            "seed": w.seed, "interp": w.interp, "upwind": w.upwind, "Re": w.Re,
            "Minf": w.Minf, "Pr": w.Pr, "gamma": w.gamma, "eps": w.weno_eps,
            "field": w.field, "time_scheme_type": w.time_scheme_type, "weno_type": w.weno_type,
            "reference_wall_s": wall, "reference_threads": 1,
            "u0_summary": summary(u0), "uN_summary": summary(uN), "trace": trace, **extra}
    arrays = {"meta": np.array(json.dumps(meta))}
    if full:
        arrays["u0"] = u0
        arrays["uN"] = uN
    else:
        arrays["uN_sub4"] = subsample(uN)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
    print(f"{name}: dims={dims} steps={steps} dt={w.dt:.6g} reference {wall:.1f}s", flush=True)


def make_ensemble(name: str):
    sizes, steps = ENSEMBLE[name]
    members = [dataclasses.replace(decks.config0(s[0], steps), dims=tuple(s), seed=20262000 + i,
                                   dt=None) for i, s in enumerate(sizes)]
    u0s = [decks.initial_state(m) for m in members]
    dt = min(decks.cfl_dt(u, m.dims) for u, m in zip(u0s, members))
    with tempfile.TemporaryDirectory() as d:
        decks.write_ensemble_deck(d, members, steps, dt)
        wall, out = run_reference(d, want_stdout=True)
        trace = step_trace(out)
        assert len(trace["cfl"]) == steps
        outs = []
        for i in range(len(members)):
            _, _, uN = decks.read_op_bin(os.path.join(d, f"op_{decks.member_index(i, len(members))}.bin"))
            outs.append(uN)
    meta = {"name": name, "sizes": [list(s) for s in sizes], "steps": steps, "dt": dt,
            "seeds": [m.seed for m in members], "reference_wall_s": wall,
            "reference_threads": 1, "trace": trace}
    arrays = {"meta": np.array(json.dumps(meta))}
    for i, (u0, uN) in enumerate(zip(u0s, outs)):
        arrays[f"u0_{i}"] = u0
        arrays[f"uN_{i}"] = uN
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
    print(f"{name}: sizes={sizes} steps={steps} dt={dt:.6g} reference {wall:.1f}s", flush=True)


def main(argv):
    if not os.path.exists(REF_BIN):
        raise SystemExit(f"{REF_BIN} missing: make -f oracle/Makefile.ref REF=/root/reference")
    names = argv[1:] or (list(FIXTURES) + list(ENSEMBLE))
    if names == ["--fast"]:  # all but the two long runs (minutes each, single-threaded)
        names = [n for n in list(FIXTURES) + list(ENSEMBLE) if n not in ("c0_64_s100", "c1_64_s200")]
    for n in names:
        if n in FIXTURES:
            make_single(n)
        elif n in ENSEMBLE:
            make_ensemble(n)
        else:
            raise SystemExit(f"unknown fixture {n}")


if __name__ == "__main__":
    main(sys.argv)
