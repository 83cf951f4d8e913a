"""The C-ABI library loads and exports every entry point the public header
declares (no device calls: runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "hypar_mi355x.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("hm_create", "hm_rk4", "hm_rhs", "hm_hyperbolic", "hm_parabolic", "hm_cfl"):
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
