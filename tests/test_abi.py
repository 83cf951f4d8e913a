"""The C-ABI library loads and exports every entry point the public headers
declare (no device calls: runs without a GPU)."""
import ctypes
import glob
import os
import re

from conftest import ROOT

HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def declared_symbols():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(hm_[a-zA-Z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("hm_create", "hm_rk4", "hm_rhs", "hm_hyperbolic", "hm_parabolic", "hm_cfl", "hm_bind",
              "hm_slot_TimeIntegrate", "hm_slot_RHSFunction", "hm_slot_HyperbolicFunction",
              "hm_slot_ParabolicFunction", "hm_slot_ComputeCFL"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from hypar_amd import lib
    L = lib.load()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    declared_in_binding = {name for name, _, _ in lib.SIGNATURES}
    assert set(declared_symbols()) <= declared_in_binding


def test_version_and_error_strings_without_gpu():
    from hypar_amd import lib
    L = lib.load()
    assert b"gfx950" in L.hm_version()
    assert isinstance(L.hm_last_error(), bytes)


def test_create_rejects_bad_config_before_touching_device():
    from hypar_amd import lib
    L = lib.load()
    cfg = lib.HMConfig()
    cfg.dim[:] = [4, 16, 16]  # too small
    cfg.gamma, cfg.Re, cfg.Minf, cfg.Pr, cfg.weno_eps = 1.4, -1.0, 1.0, 0.72, 1e-6
    h = ctypes.c_void_p()
    assert L.hm_create(ctypes.byref(cfg), ctypes.byref(h)) == lib.HM_ERR_ARG
    assert b">= 5" in L.hm_last_error()
    cfg.dim[:] = [16, 16, 16]
    cfg.interp = 7
    assert L.hm_create(ctypes.byref(cfg), ctypes.byref(h)) == lib.HM_ERR_UNSUPPORTED
    cfg.interp = 0
    cfg.weno_type = 5  # one past the optimal (no_limiting) weights
    assert L.hm_create(ctypes.byref(cfg), ctypes.byref(h)) == lib.HM_ERR_UNSUPPORTED
    assert b"WENO weight" in L.hm_last_error()


def test_noalloc_cfl_rejects_oversized_grid_before_touching_device():
    from hypar_amd import lib
    L = lib.load()
    cfg = lib.HMConfig()
    cfg.dim[:] = [16, 16, 4096]
    cfg.gamma, cfg.Re, cfg.Minf, cfg.Pr, cfg.weno_eps = 1.4, -1.0, 1.0, 0.72, 1e-6
    out = ctypes.c_double()
    code = L.hm_cfl_noalloc(ctypes.byref(cfg), ctypes.c_void_p(8), 0.01, ctypes.byref(out), None)
    assert code == lib.HM_ERR_ARG
    assert b"HM_NOALLOC_MAX_SIDE" in L.hm_last_error()


def test_slot_bindings_reject_unbound_objects_before_touching_device():
    """The slots report an unbound object, and the binding calls check their
    arguments, without a device: a stand-in solver pointer is stored and
    never dereferenced on these paths."""
    from hypar_amd import lib
    L = lib.load()
    hy_buf, ts_buf, fake = (ctypes.create_string_buffer(64) for _ in range(3))
    hy, ts = ctypes.cast(hy_buf, ctypes.c_void_p), ctypes.cast(ts_buf, ctypes.c_void_p)
    assert L.hm_slot_RHSFunction(None, None, hy, None, 0.0) == lib.HM_ERR_ARG
    assert b"not bound" in L.hm_last_error()
    assert L.hm_slot_ComputeCFL(hy, None, 0.1, 0.0) < 0
    assert L.hm_slot_TimeIntegrate(ts) != 0
    assert b"no members" in L.hm_last_error()
    assert L.hm_bind(None, ctypes.cast(fake, ctypes.c_void_p)) == lib.HM_ERR_ARG
    assert L.hm_bind_state(hy, None, lib.SLOT_ARRAYS_DEVICE) == lib.HM_ERR_ARG
    assert L.hm_bind_time(ts, hy, 0.1, lib.TIME_RK4) == lib.HM_ERR_ARG
    assert L.hm_bind(hy, ctypes.cast(fake, ctypes.c_void_p)) == lib.HM_OK
    assert L.hm_bind_state(hy, None, 7) == lib.HM_ERR_ARG
    assert L.hm_bind_time(ts, hy, 0.1, 9) == lib.HM_ERR_UNSUPPORTED
    assert L.hm_bind_time(ts, hy, 0.0, lib.TIME_RK4) == lib.HM_ERR_ARG
    assert L.hm_bind_time(ts, hy, 0.1, lib.TIME_SSPRK3) == lib.HM_OK
    # bound but without a state: TimeIntegrate and ComputeCFL say so
    assert L.hm_slot_TimeIntegrate(ts) == lib.HM_ERR_ARG
    assert b"no state" in L.hm_last_error()
    assert L.hm_slot_ComputeCFL(hy, None, 0.1, 0.0) < 0
    # LimFlag other than 1 is refused before any work
    assert L.hm_slot_HyperbolicFunction(None, None, hy, None, 0.0, 0, ctypes.c_void_p(1), None) == \
        lib.HM_ERR_UNSUPPORTED
    L.hm_unbind(hy)
    L.hm_unbind(ts)
    assert L.hm_slot_RHSFunction(None, None, hy, None, 0.0) == lib.HM_ERR_ARG
