"""The CPU oracle against golden vectors written by the reference binary.

The oracle keeps the reference's operation order, so it must reproduce the
reference bit for bit (both are gcc builds without FMA contraction).
"""
import numpy as np
import pytest

from conftest import load_fixture, workload_from_meta
from hypar_amd import decks
from oracle.oracle import Oracle

WEIGHTS = [f"{c}_{t}_16_s4" for c in ("c0", "c1") for t in ("mapped", "borges", "yc", "optimal")]
SSP = ["c0_ssprk3_16_s4", "dns_16_s4", "dns_ragged_s3"]
# forward Euler, "22" and "33" (TimeExplicitRKInitialize.c:26-57)
TABLEAUX = ["c0_rk1fe_16_s2", "c0_rk22_16_s3", "c1_rk33_16_s3"]
FULL = ["c0_16_s1", "c0_16_s10", "c1_16_s1", "c1_16_s10", "c0_ragged_s3", "c1_ragged_s3"] + WEIGHTS + SSP + TABLEAUX


@pytest.mark.parametrize("name", FULL)
def test_oracle_matches_reference_bitwise(name):
    meta, z = load_fixture(name)
    w = workload_from_meta(meta)
    u0 = z["u0"]
    # the generator reproduces the stored input exactly
    assert np.array_equal(decks.initial_state(w), u0)
    got = Oracle.for_workload(w).step(u0, w.dt, meta["steps"], w.time_scheme_type)
    assert np.array_equal(got, z["uN"]), f"max diff {np.abs(got - z['uN']).max()}"


def _check_sub(name):
    meta, z = load_fixture(name)
    w = workload_from_meta(meta)
    u0 = decks.initial_state(w)
    got = Oracle.for_workload(w).rk4(u0, w.dt, meta["steps"])
    assert np.array_equal(got[::4, ::4, ::4, :], z["uN_sub4"])
    s = got.reshape(-1, 5).sum(axis=0)
    assert np.allclose(s, meta["uN_summary"]["sum"], rtol=1e-14, atol=1e-14)


def test_oracle_c1_32_20_steps():
    _check_sub("c1_32_s20")


@pytest.mark.slow
@pytest.mark.parametrize("name", ["c0_64_s100", "c1_64_s200"])
def test_oracle_long_runs(name):
    _check_sub(name)


def test_oracle_ensemble_members_share_dt():
    """Ensemble members advance independently with sim[0]'s dt
    (TimeInitialize.c:49-52, TimeRK.c:126-195)."""
    meta, z = load_fixture("ens3_s4")
    for i, dims in enumerate(meta["sizes"]):
        o = Oracle(tuple(dims))
        got = o.rk4(z[f"u0_{i}"], meta["dt"], meta["steps"])
        assert np.array_equal(got, z[f"uN_{i}"])


@pytest.mark.parametrize("name", WEIGHTS)
def test_weight_formulation_changes_the_result(name):
    """The mapped/Z/YC/optimal fixtures differ from the Jiang-Shu result of
    the same workload, so a fixture cannot pass on JS weights."""
    import dataclasses
    meta, z = load_fixture(name)
    w = dataclasses.replace(workload_from_meta(meta), weno_type="js")
    js = Oracle.for_workload(w).rk4(z["u0"], w.dt, meta["steps"])
    assert decks.rel_l2(js, z["uN"]) > 1e-9


def test_oracle_volume_integral_matches_reference_print():
    """The oracle's VolumeIntegral against the value the reference printed
    for the same initial state (%1.16E, so 17 significant digits)."""
    for name in ["c1_16_s4_cons", "c0_ragged_s3_cons"]:
        meta, z = load_fixture(name)
        got = Oracle.for_workload(workload_from_meta(meta)).volume_integral(z["u0"])
        ref = np.array(meta["volume_integral_initial"])
        assert np.all(np.abs(got - ref) <= 1e-15 * np.maximum(np.abs(ref), 1e-3)), (name, got, ref)


# fixtures whose traces the oracle replays (every full fixture of the RK4 tableau)
TRACED = FULL + ["c1_16_s4_cons", "c0_ragged_s3_cons"]


def printed_trace(o, u0, dt, steps, scheme="44"):
    """The CFL before each step and the RMS of each step's change, formatted
    as the reference prints them (TimePrintStep.c:28-33).  The norm sums the
    squares point by point with the variables inner (ArraySumSquarenD,
    include/arrayfunctions.h), which a sequential cumsum over the AoS array
    reproduces; it returns the raw sums too, for the ensemble total."""
    import math
    u, cfl, norm, sums = u0, [], [], []
    npts = u0.size // 5
    for _ in range(steps):
        cfl.append("%1.3E" % o.cfl(u, dt))
        un = o.step(u, dt, 1, scheme)
        d = (u - un).ravel()
        sums.append(float(np.cumsum(d * d)[-1]))
        norm.append("%1.4E" % math.sqrt(sums[-1] / npts))
        u = un
    return cfl, norm, sums


@pytest.mark.parametrize("name", TRACED)
def test_oracle_cfl_and_norm_match_reference_print(name):
    """orc_cfl at every step's state and the RMS step change against the
    reference's own printed CFL= and norm= values, string for string."""
    meta, z = load_fixture(name)
    w = workload_from_meta(meta)
    cfl, norm, _ = printed_trace(Oracle.for_workload(w), z["u0"], w.dt, meta["steps"], w.time_scheme_type)
    assert cfl == meta["trace"]["cfl"]
    assert norm == meta["trace"]["norm"]


def test_oracle_ensemble_cfl_and_norm_match_reference_print():
    """Ensembles print the last member's CFL (the member loop of
    TimePreStep.c:93-97 overwrites max_cfl) and the norm over all members,
    sqrt(sum of every member's squares / sum of their points)
    (TimePostStep.c:43-63)."""
    import math
    meta, z = load_fixture("ens3_s4")
    per = []
    for i, dims in enumerate(meta["sizes"]):
        per.append(printed_trace(Oracle(tuple(dims)), z[f"u0_{i}"], meta["dt"], meta["steps"]))
    npts = sum(int(np.prod(d)) for d in meta["sizes"])
    cfl = per[-1][0]
    norm = []
    for n in range(meta["steps"]):
        total = 0.0
        for p in per:
            total += p[2][n]
        norm.append("%1.4E" % math.sqrt(total / npts))
    assert cfl == meta["trace"]["cfl"]
    assert norm == meta["trace"]["norm"]


@pytest.mark.parametrize("d,dims,v", [(0, (14, 12, 10), (-0.3, 0.2, -0.1)),
                                        (1, (10, 16, 12), (0.05, -0.4, 0.1)),
                                        (2, (12, 10, 18), (0.1, 0.15, -0.35))])
def test_oracle_cfl_closed_form(d, dims, v):
    """A uniform state whose largest |v|, negative, lies along direction d,
    the finest one: the CFL is (|v_d| + c) dt max_i dxinv_d(i) with
    c = sqrt(gamma p / rho).  A CFL without the absolute value, or with the
    directions' metrics swapped, gives another number."""
    o = Oracle(dims)
    rho, p, g = 1.25, 0.9, 1.4
    E = p / (g - 1) + 0.5 * rho * sum(c * c for c in v)
    u = np.empty(dims[::-1] + (5,))
    u[...] = [rho, rho * v[0], rho * v[1], rho * v[2], E]
    dt = 0.01
    c = np.sqrt(g * p / rho)
    dx = [decks.dxinv_1d(decks.grid_1d(n))[3:-3] for n in dims]
    want = max((abs(v[e]) + c) * dt * dx[e].max() for e in range(3))
    assert abs(o.cfl(u, dt) - want) <= 1e-14 * want
    assert abs(want - (abs(v[d]) + c) * dt * dx[d].max()) <= 1e-14 * want
