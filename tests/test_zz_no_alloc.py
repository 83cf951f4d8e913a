"""hm_cfl_noalloc (the entry point that allocates nothing) against hm_cfl,
through the C-ABI: two calls in a row on one GPU, the second on a smaller
grid, so the second reuses the static scratch the first left behind."""
import ctypes
import dataclasses

import numpy as np
import pytest

from hypar_amd import decks, lib

pytestmark = pytest.mark.gpu


def _cfg(dims, dxinv=None):
    c = lib.HMConfig()
    c.dim[:] = list(dims)
    c.gamma, c.Re, c.Minf, c.Pr, c.weno_eps = 1.4, -1.0, 1.0, 0.72, 1e-6
    if dxinv is not None:
        c.dxinv = dxinv.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    return c


def _noalloc(L, cfg, u, dt):
    import torch
    out = ctypes.c_double(-1.0)
    code = L.hm_cfl_noalloc(ctypes.byref(cfg), ctypes.c_void_p(u.data_ptr()), float(dt), ctypes.byref(out),
                            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return code, out.value


def test_noalloc_cfl_matches_hm_cfl_on_two_calls():
    import torch

    from hypar_amd.solver import NS3D
    L = lib.load()
    cases = [((40, 36, 48), None), ((16, 12, 20), "stretched")]
    for dims, metric in cases:
        print(f"noalloc cfl: grid {dims}, metrics {metric or 'uniform'}", flush=True)
        w = dataclasses.replace(decks.config0(16, 1), dims=dims, dt=None)
        u0, dt = decks.prepare(w)
        dxinv = None
        if metric:  # a non-uniform HyPar::dxinv array, x|y|z with 3 ghosts each
            rng = np.random.default_rng(7)
            dxinv = np.concatenate([decks.dxinv_1d(decks.grid_1d(n)) * (1.0 + 0.3 * rng.random(n + 6)) for n in dims])
        s = NS3D(dims, dxinv=dxinv)
        u = s.to_device(u0)
        ref = s.cfl(u, dt)
        code, got = _noalloc(L, _cfg(dims, dxinv), u, dt)
        assert code == lib.HM_OK, L.hm_last_error()
        assert got == ref, (dims, got, ref)  # a max: bit-exact
        s.close()
    torch.cuda.synchronize()


def test_noalloc_cfl_flags_nonfinite_and_rejects_big_grids():
    import torch
    L = lib.load()
    dims = (16, 12, 20)
    u = torch.ones(5 * 16 * 12 * 20, dtype=torch.float64, device="cuda")
    u[4 * 16 * 12 * 20 + 77] = float("nan")
    code, _ = _noalloc(L, _cfg(dims), u, 0.01)
    assert code == lib.HM_ERR_NONFINITE
    code, _ = _noalloc(L, _cfg((16, 12, 4096)), u, 0.01)
    assert code == lib.HM_ERR_ARG and b"HM_NOALLOC_MAX_SIDE" in L.hm_last_error()
