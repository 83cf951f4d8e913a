"""The C host layer (hypar_amd/hypar_mi355x): reference input decks in,
reference binary solutions out.

CPU tests check that decks the GPU path does not implement are rejected
with a clear message before any device work; GPU tests run decks written
from the golden fixtures and compare op.bin with the reference's output.
"""
import dataclasses
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_fixture, workload_from_meta
from hypar_amd import decks

DRIVER = os.path.join(ROOT, "hypar_amd", "hypar_mi355x")


def run_driver(d, *args, timeout=300):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} missing: run __graft_entry__.build()")
    return subprocess.run([DRIVER, *args, d], capture_output=True, text=True, timeout=timeout)


def _deck(tmp_path, name="c0_16_s1", conservation_check="no"):
    meta, z = load_fixture(name)
    w = workload_from_meta(meta)
    d = str(tmp_path / name)
    decks.write_deck(d, w, u0=z["u0"], screen_op_iter=1, conservation_check=conservation_check)
    return d, w, meta, z


def _edit(path, old, new):
    with open(path) as f:
        text = f.read()
    assert old in text
    with open(path, "w") as f:
        f.write(text.replace(old, new))


# ------------------------------------------------------------------ CPU
def test_rejects_gravity(tmp_path):
    d, *_ = _deck(tmp_path)
    _edit(os.path.join(d, "physics.inp"), "end", "  gravity 0.0 0.0 -9.8\nend")
    r = run_driver(d)
    assert r.returncode != 0 and "gravity is not supported" in r.stderr


def test_rejects_non_periodic_boundary(tmp_path):
    d, *_ = _deck(tmp_path)
    _edit(os.path.join(d, "boundary.inp"), "periodic 2 -1", "extrapolate 2 -1")
    r = run_driver(d)
    assert r.returncode != 0 and "periodic" in r.stderr


def test_rejects_domain_decomposition(tmp_path):
    d, *_ = _deck(tmp_path)
    _edit(os.path.join(d, "solver.inp"), "iproc               1 1 1", "iproc               2 1 1")
    r = run_driver(d)
    assert r.returncode != 0 and "iproc" in r.stderr


def test_missing_solver_inp(tmp_path):
    r = run_driver(str(tmp_path))
    assert r.returncode != 0 and "solver.inp" in r.stderr


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
@pytest.mark.parametrize("name,args", [("c0_16_s10", []), ("c1_16_s10", []), ("c1_16_s10", ["--slots"]),
                                       ("c0_16_s1", ["--slots"]), ("c0_ssprk3_16_s4", []),
                                       ("c0_ssprk3_16_s4", ["--slots"]), ("dns_16_s4", []), ("dns_ragged_s3", ["--slots"]),
                                       ("c1_mapped_16_s4", []), ("c0_yc_16_s4", ["--slots"]),
                                       ("c1_borges_16_s4", ["--slots"]), ("c1_optimal_16_s4", []),
                                       ("c0_optimal_16_s4", ["--slots"]), ("c0_rk1fe_16_s2", []),
                                       ("c0_rk22_16_s3", []), ("c1_rk33_16_s3", []),
                                       ("c1_rk33_16_s3", ["--slots"])])
def test_driver_matches_reference(tmp_path, name, args):
    d, w, meta, z = _deck(tmp_path, name)
    print(f"[gpu] driver {name} {args}", flush=True)
    r = run_driver(d, *args)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    scheme = meta.get("time_scheme_type", "44")
    if args:
        assert "slot composition" in r.stdout
    elif scheme == "ssprk3":
        assert "fused SSPRK3 step" in r.stdout
    elif scheme == "44":
        assert "fused 3N RK4 step" in r.stdout
    else:  # HyPar's own stage loop with the library's RHSFunction slot
        assert "stage loop around the fused right-hand side" in r.stdout
    _, _, u = decks.read_op_bin(os.path.join(d, "op.bin"))
    assert decks.rel_l2(u, z["uN"]) <= 1e-12
    # function_counts.dat (SimulationWriteErrors.c:91-95): one hyperbolic,
    # parabolic (viscous only) and source evaluation per stage
    # (HyperbolicFunction.c:67, NavierStokes3DParabolicFunction.c:71-73,
    # SourceFunction.c:46-49)
    nrhs = meta["steps"] * {"44": 4, "ssprk3": 3, "33": 3, "22": 2, "1fe": 1}[scheme]
    with open(os.path.join(d, "function_counts.dat")) as f:
        counts = [int(x) for x in f.read().split()]
    assert counts == [meta["steps"], nrhs, nrhs if meta["Re"] > 0 else 0, nrhs], counts


@pytest.mark.gpu
@pytest.mark.parametrize("args", [[], ["--slots"]])
def test_driver_ensemble_matches_reference(tmp_path, args):
    meta, z = load_fixture("ens3_s4")
    members = [dataclasses.replace(decks.config0(s[0], meta["steps"]), dims=tuple(s), seed=sd, dt=None)
               for s, sd in zip(meta["sizes"], meta["seeds"])]
    d = str(tmp_path / "ens")
    decks.write_ensemble_deck(d, members, meta["steps"], meta["dt"])
    print(f"[gpu] driver ensemble {args}", flush=True)
    r = run_driver(d, *args)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    # fused: one launch per pass for all members; --slots: member by member
    assert ("batched over all members" in r.stdout) == (not args)
    for i in range(len(members)):
        _, _, u = decks.read_op_bin(os.path.join(d, f"op_{i}.bin"))
        assert decks.rel_l2(u, z[f"uN_{i}"]) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name,args", [("c1_16_s4_cons", []), ("c1_16_s4_cons", ["--slots"]),
                                       ("c0_ragged_s3_cons", [])])
def test_driver_conservation_check_matches_reference(tmp_path, name, args):
    """conservation_check yes: the printed initial volume integral, and
    conservation.dat, against the reference's for the same deck.  The viscous
    deck's momentum and energy errors are O(1e-8): the reference's ghost-
    layer viscous fluxes do not telescope, and this path evaluates them the
    same way.  The inviscid deck's errors are rounding noise (~1e-15)."""
    d, w, meta, z = _deck(tmp_path, name, conservation_check="yes")
    print(f"[gpu] driver conservation {name} {args}", flush=True)
    r = run_driver(d, *args)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    lines = r.stdout.splitlines()
    i = next(k for k, l in enumerate(lines) if l.startswith("Volume integral of the initial solution"))
    vi = np.array([float(lines[i + 1 + v].split()[1]) for v in range(5)])
    ref = np.array(meta["volume_integral_initial"])
    scale = np.maximum(np.abs(ref), 1.0)
    assert np.all(np.abs(vi - ref) <= 1e-13 * scale), (vi, ref)
    with open(os.path.join(d, "conservation.dat")) as f:
        fields = f.read().split()
    rf = meta["conservation_dat"]
    assert fields[:7] == rf[:7]  # sizes, processes, dt
    got, want = np.array([float(x) for x in fields[7:]]), np.array([float(x) for x in rf[7:]])
    print(f"[gpu]   conservation errors {got} reference {want}", flush=True)
    assert np.all(np.abs(got - want) <= 1e-12 + 1e-4 * np.abs(want)), (got, want)
    assert "Conservation Errors:" in r.stdout
    _, _, u = decks.read_op_bin(os.path.join(d, "op.bin"))
    assert decks.rel_l2(u, z["uN"]) <= 1e-12
