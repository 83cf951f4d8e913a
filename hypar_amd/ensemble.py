"""Batched ensembles of NS3D grids through the C-ABI (hm_ensemble_*).

Members may have uneven grid sizes and each has its own dt; they share the
scheme, inviscid or viscous.  The batched device state is one float64 tensor
holding member m's hm-layout state at [5*offset_m, 5*(offset_m + ncell_m)).
Work runs on torch's current stream.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np
import torch

from . import lib as _lib


class Ensemble:
    def __init__(self, dims: Sequence[Sequence[int]], interp: str = "components", upwind: str = "rusanov",
                 gamma: float = 1.4, eps: float = 1e-6, Re: float = -1.0, Minf: float = 1.0, Pr: float = 0.72,
                 weno_type: str = "js"):
        if not torch.cuda.is_available():
            raise RuntimeError("Ensemble needs a GPU (torch.cuda.is_available() is False)")
        self.L = _lib.load()
        self.dims = [tuple(int(d) for d in m) for m in dims]
        n = len(self.dims)
        cfgs = (_lib.HMConfig * n)()
        for c, d in zip(cfgs, self.dims):
            c.dim[:] = list(d)
            c.gamma, c.Re, c.Minf, c.Pr = gamma, Re, Minf, Pr
            c.interp = _lib.INTERP[interp]
            c.upwind = _lib.UPWIND[upwind]
            c.weno_eps = eps
            c.weno_type = _lib.WENO[weno_type]
        torch.cuda.init()
        h = ctypes.c_void_p()
        _lib.check(self.L.hm_ensemble_create(cfgs, n, ctypes.byref(h)))
        self.h = h
        self.ncell = int(self.L.hm_ensemble_ncell(self.h))
        self.offsets, self.sizes = [], []
        for m in range(n):
            off, nc = ctypes.c_size_t(), ctypes.c_size_t()
            _lib.check(self.L.hm_ensemble_member(self.h, m, ctypes.byref(off), ctypes.byref(nc)))
            self.offsets.append(off.value)
            self.sizes.append(nc.value)

    def close(self):
        if getattr(self, "h", None):
            self.L.hm_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self):
        _lib.check(self.L.hm_ensemble_set_stream(self.h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def member(self, u: torch.Tensor, m: int) -> torch.Tensor:
        """View of member m's hm-layout state inside the batched state."""
        o, n = self.offsets[m], self.sizes[m]
        return u[5 * o:5 * (o + n)]

    def to_device(self, states) -> torch.Tensor:
        """Batched device state from per-member (nz, ny, nx, 5) host arrays."""
        u = torch.empty(5 * self.ncell, dtype=torch.float64, device="cuda")
        for m, a in enumerate(states):
            t = torch.as_tensor(np.ascontiguousarray(a)).permute(3, 0, 1, 2).reshape(-1)
            self.member(u, m).copy_(t)
        return u

    def to_host(self, u: torch.Tensor, m: int) -> np.ndarray:
        nx, ny, nz = self.dims[m]
        return self.member(u, m).view(5, nz, ny, nx).permute(1, 2, 3, 0).contiguous().cpu().numpy()

    def _check(self, u):
        if not (u.is_cuda and u.dtype == torch.float64 and u.is_contiguous() and u.numel() == 5 * self.ncell):
            raise ValueError("batched state must be a contiguous float64 CUDA tensor of 5*ncell entries")

    def set_dt(self, dt: Sequence[float]):
        self._stream()
        arr = (ctypes.c_double * len(self.dims))(*[float(x) for x in dt])
        _lib.check(self.L.hm_ensemble_set_dt(self.h, arr))

    def rk4(self, u: torch.Tensor, nsteps: int) -> torch.Tensor:
        self._check(u)
        self._stream()
        _lib.check(self.L.hm_ensemble_rk4(self.h, ctypes.c_void_p(u.data_ptr()), int(nsteps)))
        return u

    # -- batched discrete adjoint ---------------------------------------
    def adjoint_forward(self, u: torch.Tensor, nsteps: int, every: int = 10) -> torch.Tensor:
        """Advance u nsteps RK4 steps in place, checkpointing every `every` steps."""
        self._check(u)
        self._stream()
        _lib.check(self.L.hm_ensemble_adjoint_forward(self.h, ctypes.c_void_p(u.data_ptr()), int(nsteps), int(every)))
        self.tape_id = getattr(self, "tape_id", 0) + 1
        return u

    def adjoint_reverse(self, lam: torch.Tensor) -> torch.Tensor:
        """lam: dJ/du_N of all members in, dJ/du_0 out (in place)."""
        self._check(lam)
        self._stream()
        _lib.check(self.L.hm_ensemble_adjoint_reverse(self.h, ctypes.c_void_p(lam.data_ptr())))
        return lam

    def set_tiles(self, shape: int, grid=(0, 0, 0)):
        g = (ctypes.c_uint * 3)(*grid)
        _lib.check(self.L.hm_ensemble_set_tiles(self.h, int(shape), g))

    def set_tape_budget(self, nbytes: int):
        """Bytes the batched adjoint tape may hold (0: from the device's free memory)."""
        _lib.check(self.L.hm_ensemble_set_tape_budget(self.h, int(nbytes)))

    def tape_info(self) -> dict:
        """Schedule of the last adjoint_forward: sets, step level, bytes."""
        sets, step, nb = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
        _lib.check(self.L.hm_ensemble_tape_info(self.h, ctypes.byref(sets), ctypes.byref(step), ctypes.byref(nb)))
        return {"sets": sets.value, "step_level": bool(step.value), "bytes": nb.value}

    def set_wait_limit(self, polls: int):
        """Polls after which a dependency wait of the one-launch-per-stage
        kernel (shapes 2, 3) gives up and the call reports HM_ERR_TIMEOUT."""
        _lib.check(self.L.hm_ensemble_set_wait_limit(self.h, int(polls)))

    def tiles(self):
        sh, g = ctypes.c_int(), (ctypes.c_uint * 3)()
        _lib.check(self.L.hm_ensemble_get_tiles(self.h, ctypes.byref(sh), g))
        return sh.value, tuple(g)

    def tune(self, u: torch.Tensor, reps: int = 2):
        """Autotune tile shape and persistent grid size on a copy of u."""
        self._check(u)
        self._stream()
        sh, g = ctypes.c_int(), (ctypes.c_uint * 3)()
        _lib.check(self.L.hm_ensemble_tune(self.h, ctypes.c_void_p(u.data_ptr()), int(reps), ctypes.byref(sh), g))
        return sh.value, tuple(g)
