"""HyPar's slots filled with the library's exports (include/hypar_mi355x_slots.h).

A C caller holds each operator as a function pointer of the reference's own
type (include/hypar.h:220-276, include/timeintegration_struct.h:96-98 of the
reference) and passes its solver and time-integration objects as opaque
pointers.  These tests do the same through ctypes: the hm_slot_* addresses
go into structs of CFUNCTYPE fields, opaque buffers stand for HyPar's
objects, and every call goes through the stored pointer.  Results are
checked against the reference's fixtures (final states, and the CFL numbers
it printed before each step) and against the library's direct entry points.
"""
import ctypes

import numpy as np
import pytest

from conftest import load_fixture, within_print as _within_print, workload_from_meta

torch = pytest.importorskip("torch")
from hypar_amd import lib  # noqa: E402  (after torch: one HIP runtime per process)

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(msg):
    print(msg, flush=True)


class SolverSlots(ctypes.Structure):
    """The slots of HyPar's solver object that this path fills."""
    _fields_ = [("HyperbolicFunction", lib.HyperbolicFunctionProto),
                ("ParabolicFunction", lib.ParabolicFunctionProto),
                ("ComputeCFL", lib.ComputeCFLProto)]


class TimeSlots(ctypes.Structure):
    """TimeIntegration's two slots."""
    _fields_ = [("TimeIntegrate", lib.TimeIntegrateProto), ("RHSFunction", lib.RHSFunctionProto)]


def _install():
    s = SolverSlots()
    s.HyperbolicFunction = lib.slot_pointer("hm_slot_HyperbolicFunction", lib.HyperbolicFunctionProto)
    s.ParabolicFunction = lib.slot_pointer("hm_slot_ParabolicFunction", lib.ParabolicFunctionProto)
    s.ComputeCFL = lib.slot_pointer("hm_slot_ComputeCFL", lib.ComputeCFLProto)
    t = TimeSlots()
    t.TimeIntegrate = lib.slot_pointer("hm_slot_TimeIntegrate", lib.TimeIntegrateProto)
    t.RHSFunction = lib.slot_pointer("hm_slot_RHSFunction", lib.RHSFunctionProto)
    return s, t


def _opaque():
    """A stand-in for one of HyPar's objects: the slots only use its address."""
    buf = ctypes.create_string_buffer(512)
    return buf, ctypes.cast(buf, ctypes.c_void_p)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _flux_callbacks(calls):
    """Stand-ins for HyPar's FFunction and Upwind pointers: the library's
    slot must not call them."""

    def f(*a):
        calls.append("FFunction")
        return 1

    def up(*a):
        calls.append("Upwind")
        return 1
    return lib.FFunctionProto(f), lib.UpwindProto(up)


def test_slots_on_device_arrays_match_reference_fixture():
    """Viscous characteristic WENO5 + Roe, RK4, 16^3 x 10 steps: the state
    advanced by TimeIntegrate and the CFL number before every step."""
    from hypar_amd import decks
    from hypar_amd.solver import NS3D
    L = lib.load()
    meta, z = load_fixture("c1_16_s10")
    w = workload_from_meta(meta)
    _say(f"[gpu] slot ABI, device arrays: {meta['name']} dims={w.dims}")
    s = NS3D.from_workload(w)
    u = s.to_device(z["u0"])
    hy_buf, hy = _opaque()
    ts_buf, ts = _opaque()
    lib.check(L.hm_bind(hy, s.h))
    lib.check(L.hm_bind_state(hy, _ptr(u), lib.SLOT_ARRAYS_DEVICE))
    lib.check(L.hm_bind_time(ts, hy, w.dt, lib.TIME_RK4))
    S, T = _install()
    for n in range(meta["steps"]):
        cfl = S.ComputeCFL(hy, None, w.dt, n * w.dt)
        assert _within_print(cfl, meta["trace"]["cfl"][n], 3), (n, cfl, meta["trace"]["cfl"][n])
        assert T.TimeIntegrate(ts) == 0, L.hm_last_error()
    err = decks.rel_l2(s.to_host(u), z["uN"])
    _say(f"[gpu]   TimeIntegrate x {meta['steps']}: rel L2 vs reference = {err:.3e}")
    assert err <= TOL

    # the term slots equal the library's direct entry points bit for bit,
    # and RHSFunction is -hyperbolic + parabolic (TimeRHSFunctionExplicit.c:88-93)
    calls = []
    F, U = _flux_callbacks(calls)
    rhs, hyp, par = s.empty(), s.empty(), s.empty()
    assert T.RHSFunction(_ptr(rhs), _ptr(u), hy, None, 0.0) == 0
    assert S.HyperbolicFunction(_ptr(hyp), _ptr(u), hy, None, 0.0, 1, F, U) == 0
    assert S.ParabolicFunction(_ptr(par), _ptr(u), hy, None, 0.0) == 0
    torch.cuda.synchronize()
    assert not calls
    assert torch.equal(rhs, s.rhs(u))
    assert torch.equal(hyp, s.hyperbolic(u))
    assert torch.equal(par, s.parabolic(u))
    assert torch.equal(rhs, (-hyp) + par)

    # a NULL FFunction gives a zero hyperbolic term (HyperbolicFunction.c:65-67);
    # LimFlag 0 asks for weights the fused operator does not keep
    assert S.HyperbolicFunction(_ptr(hyp), _ptr(u), hy, None, 0.0, 1, lib.FFunctionProto(), U) == 0
    torch.cuda.synchronize()
    assert int(torch.count_nonzero(hyp)) == 0
    assert S.HyperbolicFunction(_ptr(hyp), _ptr(u), hy, None, 0.0, 0, F, U) == lib.HM_ERR_UNSUPPORTED
    assert b"LimFlag" in L.hm_last_error()

    # hm_destroy drops the binding: the slot then reports an unbound object
    s.close()
    assert T.RHSFunction(_ptr(rhs), _ptr(u), hy, None, 0.0) == lib.HM_ERR_ARG
    assert b"not bound" in L.hm_last_error()
    L.hm_unbind(ts)
    del hy_buf, ts_buf


def test_slots_on_reference_host_arrays_match_reference_fixture():
    """The published DNS deck's scheme (viscous, mapped component-wise
    WENO5 + Roe, SSPRK3), 16^3 x 4 steps, with HyPar's own host arrays
    (AoS, 3 ghost layers) as the slots' arrays."""
    from hypar_amd import decks
    from hypar_amd.solver import NS3D
    L = lib.load()
    meta, z = load_fixture("dns_16_s4")
    w = workload_from_meta(meta)
    _say(f"[gpu] slot ABI, reference host arrays: {meta['name']} dims={w.dims}")
    nx, ny, nz = w.dims
    g = 3
    G = np.zeros((nz + 2 * g, ny + 2 * g, nx + 2 * g, 5))
    G[g:-g, g:-g, g:-g] = z["u0"]
    s = NS3D.from_workload(w)
    hy_buf, hy = _opaque()
    ts_buf, ts = _opaque()
    lib.check(L.hm_bind(hy, s.h))
    lib.check(L.hm_bind_state(hy, ctypes.c_void_p(G.ctypes.data), lib.SLOT_ARRAYS_REFERENCE_HOST))
    lib.check(L.hm_bind_time(ts, hy, w.dt, lib.TIME_SSPRK3))
    S, T = _install()

    # RHSFunction on host arrays equals the device-array result of the same state
    rhs_h = np.full_like(G, 7.0)
    assert T.RHSFunction(ctypes.c_void_p(rhs_h.ctypes.data), ctypes.c_void_p(G.ctypes.data), hy, None, 0.0) == 0
    ud = s.to_device(z["u0"])
    rhs_d = s.to_host(s.rhs(ud))
    assert np.array_equal(rhs_h[g:-g, g:-g, g:-g], rhs_d)
    ghost = np.ones(G.shape[:3], dtype=bool)
    ghost[g:-g, g:-g, g:-g] = False
    assert not rhs_h[ghost].any()

    for n in range(meta["steps"]):
        cfl = S.ComputeCFL(hy, None, w.dt, n * w.dt)
        assert _within_print(cfl, meta["trace"]["cfl"][n], 3), (n, cfl, meta["trace"]["cfl"][n])
        assert T.TimeIntegrate(ts) == 0, L.hm_last_error()
    err = decks.rel_l2(G[g:-g, g:-g, g:-g], z["uN"])
    _say(f"[gpu]   TimeIntegrate x {meta['steps']} on host arrays: rel L2 vs reference = {err:.3e}")
    assert err <= TOL
    L.hm_unbind(hy)
    L.hm_unbind(ts)
    s.close()
    del hy_buf, ts_buf


def test_time_integrate_advances_every_member():
    """Two solver objects with their own grids and dt on one time-integration
    object: one TimeIntegrate call steps both (the member loop of
    TimeRK.c:130-193), each matching its reference fixture."""
    from hypar_amd import decks
    from hypar_amd.solver import NS3D
    L = lib.load()
    names = ["c0_ragged_s3", "c1_ragged_s3"]
    fx = [load_fixture(n) for n in names]
    ws = [workload_from_meta(m) for m, _ in fx]
    assert fx[0][0]["steps"] == fx[1][0]["steps"]
    _say(f"[gpu] slot ABI, two members on one TimeIntegrate: {names}")
    solvers = [NS3D.from_workload(w) for w in ws]
    states = [s.to_device(z["u0"]) for s, (_, z) in zip(solvers, fx)]
    objs = [_opaque() for _ in ws]
    ts_buf, ts = _opaque()
    for (_, o), s, u, w in zip(objs, solvers, states, ws):
        lib.check(L.hm_bind(o, s.h))
        lib.check(L.hm_bind_state(o, _ptr(u), lib.SLOT_ARRAYS_DEVICE))
        lib.check(L.hm_bind_time(ts, o, w.dt, lib.TIME_RK4))
    _, T = _install()
    for _ in range(fx[0][0]["steps"]):
        assert T.TimeIntegrate(ts) == 0, L.hm_last_error()
    for name, s, u, (_, z) in zip(names, solvers, states, fx):
        err = decks.rel_l2(s.to_host(u), z["uN"])
        _say(f"[gpu]   {name}: rel L2 vs reference = {err:.3e}")
        assert err <= TOL
    L.hm_unbind(ts)
    for s in solvers:
        s.close()
    del objs, ts_buf
