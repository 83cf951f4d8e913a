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
FULL = ["c0_16_s1", "c0_16_s10", "c1_16_s1", "c1_16_s10", "c0_ragged_s3", "c1_ragged_s3"] + WEIGHTS


@pytest.mark.parametrize("name", FULL)
def test_oracle_matches_reference_bitwise(name):
    meta, z = load_fixture(name)
    w = workload_from_meta(meta)
    u0 = z["u0"]
    # the generator reproduces the stored input exactly
    assert np.array_equal(decks.initial_state(w), u0)
    got = Oracle.for_workload(w).rk4(u0, w.dt, meta["steps"])
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
