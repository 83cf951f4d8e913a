"""Python layer over the C-ABI: an NS3D solver whose state lives in HBM as a
torch float64 tensor in hm layout (SoA, no ghosts; include/hypar_mi355x.h).

Mirrors the reference's solver object for this path: a grid with its
physics and numerics fixed at construction (NavierStokes3DInitialize,
WENOInitialize), then TimeRK steps (TimeIntegrate), the right-hand side
(RHSFunction), its hyperbolic and parabolic parts, the CFL number
(ComputeCFL) and the per-step norm (TimePostStep).
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np
import torch

from . import lib as _lib


class NS3D:
    """Navier-Stokes 3D WENO5 + RK4 on the current GPU."""

    def __init__(self, dims: Sequence[int], interp: str = "components", upwind: str = "rusanov",
                 Re: float = -1.0, Minf: float = 1.0, Pr: float = 0.72, gamma: float = 1.4,
                 eps: float = 1e-6, dxinv: Optional[np.ndarray] = None, weno_type: str = "js"):
        if not torch.cuda.is_available():
            raise RuntimeError("NS3D needs a GPU (torch.cuda.is_available() is False)")
        self.L = _lib.load()
        self.dims = tuple(int(d) for d in dims)
        self.ncell = int(np.prod(self.dims))
        cfg = _lib.HMConfig()
        cfg.dim[:] = list(self.dims)
        cfg.gamma, cfg.Re, cfg.Minf, cfg.Pr = gamma, Re, Minf, Pr
        cfg.interp = _lib.INTERP[interp]
        cfg.upwind = _lib.UPWIND[upwind]
        cfg.weno_eps = eps
        cfg.weno_type = _lib.WENO[weno_type]
        self._dx = None
        if dxinv is not None:
            self._dx = np.ascontiguousarray(dxinv, dtype=np.float64)
            cfg.dxinv = self._dx.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        torch.cuda.init()
        h = ctypes.c_void_p()
        _lib.check(self.L.hm_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.tape_id = 0  # generation of the adjoint tape (one per solver)

    @classmethod
    def from_workload(cls, w):
        return cls(w.dims, w.interp, w.upwind, w.Re, w.Minf, w.Pr, w.gamma, w.weno_eps, weno_type=w.weno_type)

    # ------------------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.L.hm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self):
        _lib.check(self.L.hm_set_stream(self.h, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def _check_state(self, t: torch.Tensor, name="u"):
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
                and t.numel() == 5 * self.ncell):
            raise ValueError(f"{name} must be a contiguous float64 CUDA tensor of {5 * self.ncell} elements")

    def empty(self) -> torch.Tensor:
        return torch.empty(5 * self.ncell, dtype=torch.float64, device="cuda")

    # ------------------------------------------------------------------
    def to_device(self, u: np.ndarray, layout: str = "aos") -> torch.Tensor:
        """Upload a host state: (nz,ny,nx,5) interior AoS, or the reference's
        ghosted AoS array (layout='reference')."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        t = self.empty()
        self._stream()
        code = {"aos": _lib.LAYOUT_AOS_INTERIOR, "reference": _lib.LAYOUT_REFERENCE,
                "soa": _lib.LAYOUT_SOA_INTERIOR}[layout]
        _lib.check(self.L.hm_upload(self.h, ctypes.c_void_p(t.data_ptr()),
                                    ctypes.c_void_p(u.ctypes.data), code))
        return t

    def to_host(self, t: torch.Tensor) -> np.ndarray:
        """Download to a (nz,ny,nx,5) interior AoS array."""
        self._check_state(t)
        nx, ny, nz = self.dims
        out = np.empty((nz, ny, nx, 5), dtype=np.float64)
        self._stream()
        _lib.check(self.L.hm_download(self.h, ctypes.c_void_p(out.ctypes.data),
                                      ctypes.c_void_p(t.data_ptr()), _lib.LAYOUT_AOS_INTERIOR))
        return out

    # ------------------------------------------------------------------
    def rk4(self, u: torch.Tensor, dt: float, nsteps: int) -> torch.Tensor:
        """nsteps RK4 steps in place (enqueued on torch's current stream)."""
        self._check_state(u)
        self._stream()
        _lib.check(self.L.hm_rk4(self.h, ctypes.c_void_p(u.data_ptr()), float(dt), int(nsteps)))
        return u

    def ssprk3(self, u: torch.Tensor, dt: float, nsteps: int) -> torch.Tensor:
        """nsteps SSPRK3 steps in place (enqueued on torch's current stream)."""
        self._check_state(u)
        self._stream()
        _lib.check(self.L.hm_ssprk3(self.h, ctypes.c_void_p(u.data_ptr()), float(dt), int(nsteps)))
        return u

    def step(self, u: torch.Tensor, dt: float, nsteps: int, scheme: str = "44") -> torch.Tensor:
        """nsteps steps of the fused explicit Runge-Kutta scheme named as in
        solver.inp's time_scheme_type ("44" or "ssprk3"/"tvdrk3")."""
        if scheme == "44":
            return self.rk4(u, dt, nsteps)
        if scheme in ("ssprk3", "tvdrk3"):
            return self.ssprk3(u, dt, nsteps)
        raise ValueError(f"no fused step for time_scheme_type {scheme}")

    def _term(self, fn, u):
        self._check_state(u)
        out = self.empty()
        self._stream()
        _lib.check(fn(self.h, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(u.data_ptr())))
        return out

    def rhs(self, u):
        return self._term(self.L.hm_rhs, u)

    def hyperbolic(self, u):
        return self._term(self.L.hm_hyperbolic, u)

    def parabolic(self, u):
        return self._term(self.L.hm_parabolic, u)

    def cfl(self, u: torch.Tensor, dt: float) -> float:
        self._check_state(u)
        self._stream()
        out = ctypes.c_double()
        _lib.check(self.L.hm_cfl(self.h, ctypes.c_void_p(u.data_ptr()), float(dt), ctypes.byref(out)))
        return out.value

    def step_norm(self, a: torch.Tensor, b: torch.Tensor) -> float:
        self._check_state(a, "a")
        self._check_state(b, "b")
        self._stream()
        out = ctypes.c_double()
        _lib.check(self.L.hm_step_norm(self.h, ctypes.c_void_p(a.data_ptr()),
                                       ctypes.c_void_p(b.data_ptr()), ctypes.byref(out)))
        return out.value


    # ------------------------------------------------------------------
    # discrete adjoint (inviscid component-wise WENO5 + Rusanov)
    def volume_integral(self, u: torch.Tensor) -> np.ndarray:
        """Sum over cells of u dV for each conserved variable (synchronises)."""
        self._check_state(u)
        self._stream()
        out = (ctypes.c_double * 5)()
        _lib.check(self.L.hm_volume_integral(self.h, ctypes.c_void_p(u.data_ptr()), out))
        return np.array(out[:])

    def rhs_vjp(self, u: torch.Tensor, lam: torch.Tensor) -> torch.Tensor:
        """(d rhs/d u)^T lam at state u."""
        self._check_state(u)
        self._check_state(lam, "lam")
        out = self.empty()
        self._stream()
        _lib.check(self.L.hm_rhs_vjp(self.h, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(u.data_ptr()),
                                     ctypes.c_void_p(lam.data_ptr())))
        return out

    def rk4_vjp(self, u_n: torch.Tensor, dt: float, lam: torch.Tensor) -> torch.Tensor:
        """lam <- (d u_{n+1}/d u_n)^T lam for one RK4 step (in place)."""
        self._check_state(u_n)
        self._check_state(lam, "lam")
        self._stream()
        _lib.check(self.L.hm_rk4_vjp(self.h, ctypes.c_void_p(u_n.data_ptr()), float(dt),
                                     ctypes.c_void_p(lam.data_ptr())))
        return lam

    def adjoint_forward(self, u: torch.Tensor, dt: float, nsteps: int, every: int) -> torch.Tensor:
        """nsteps RK4 steps in place, checkpointing every `every` steps."""
        self._check_state(u)
        self._stream()
        self.tape_id += 1
        _lib.check(self.L.hm_adjoint_forward(self.h, ctypes.c_void_p(u.data_ptr()), float(dt), int(nsteps),
                                             int(every)))
        return u

    def adjoint_reverse(self, lam: torch.Tensor) -> torch.Tensor:
        """dJ/du_N -> dJ/du_0 through the last adjoint_forward (in place)."""
        self._check_state(lam, "lam")
        self._stream()
        _lib.check(self.L.hm_adjoint_reverse(self.h, ctypes.c_void_p(lam.data_ptr())))
        return lam

    def set_tape_budget(self, nbytes: int):
        """Bytes the adjoint tape may hold (0: from the device's free memory)."""
        _lib.check(self.L.hm_set_tape_budget(self.h, int(nbytes)))

    def tape_info(self) -> dict:
        """Schedule of the last adjoint_forward: sets, step level, bytes."""
        sets, step, nb = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
        _lib.check(self.L.hm_tape_info(self.h, ctypes.byref(sets), ctypes.byref(step), ctypes.byref(nb)))
        return {"sets": sets.value, "step_level": bool(step.value), "bytes": nb.value}

    def profile_kernels(self, u: torch.Tensor, reps: int = 5) -> dict:
        """Mean duration (ms) of each kernel of one RK stage, HIP events on
        the solver's stream."""
        self._check_state(u)
        self._stream()
        ms = (ctypes.c_double * 5)()
        _lib.check(self.L.hm_profile_kernels(self.h, ctypes.c_void_p(u.data_ptr()), int(reps), ms))
        return {"pass_x": ms[0], "pass_y": ms[1], "pass_z": ms[2], "visc_par": ms[3], "prims": ms[4]}


def soa_to_aos(t: torch.Tensor, dims) -> np.ndarray:
    """hm-layout tensor -> (nz,ny,nx,5) numpy (host copy through torch)."""
    nx, ny, nz = dims
    return t.view(5, nz, ny, nx).permute(1, 2, 3, 0).contiguous().cpu().numpy()
