"""ctypes loader for the C restatement oracle (libns3d_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the reference to compare against, never by the
product path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

INTERP = {"components": 0, "characteristic": 1}
UPWIND = {"roe": 0, "rusanov": 1}


class OrcParams(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int * 3), ("ghosts", ctypes.c_int),
                ("gamma", ctypes.c_double), ("Re", ctypes.c_double),
                ("Pr", ctypes.c_double), ("interp", ctypes.c_int),
                ("upwind", ctypes.c_int), ("eps", ctypes.c_double),
                ("dxinv", ctypes.POINTER(ctypes.c_double)), ("weno_type", ctypes.c_int)]

WENO_TYPES = {"js": 0, "mapped": 1, "borges": 2, "yc": 3, "optimal": 4}


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libns3d_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        pp = ctypes.POINTER(OrcParams)
        L.orc_size.argtypes = [pp]; L.orc_size.restype = ctypes.c_long
        L.orc_rhs.argtypes = [pp, dp, dp]; L.orc_rhs.restype = ctypes.c_int
        L.orc_hyperbolic.argtypes = [pp, dp, dp]; L.orc_hyperbolic.restype = ctypes.c_int
        L.orc_parabolic.argtypes = [pp, dp, dp]; L.orc_parabolic.restype = ctypes.c_int
        L.orc_rk4.argtypes = [pp, dp, ctypes.c_double, ctypes.c_int]; L.orc_rk4.restype = ctypes.c_int
        L.orc_ssprk3.argtypes = [pp, dp, ctypes.c_double, ctypes.c_int]; L.orc_ssprk3.restype = ctypes.c_int
        L.orc_rk_tableau.argtypes = [pp, dp, ctypes.c_double, ctypes.c_int, ctypes.c_int, dp, dp]
        L.orc_rk_tableau.restype = ctypes.c_int
        L.orc_cfl.argtypes = [pp, dp, ctypes.c_double]; L.orc_cfl.restype = ctypes.c_double
        L.orc_volume_integral.argtypes = [pp, dp, dp]; L.orc_volume_integral.restype = None
        L.orc_periodic_bc.argtypes = [pp, dp]; L.orc_periodic_bc.restype = None
        L.orc_weno_js_weights.argtypes = [ctypes.c_double] * 6 + [dp]
        L.orc_weno_weights.argtypes = [ctypes.c_int] + [ctypes.c_double] * 6 + [dp]
        L.orc_roe_average.argtypes = [dp, dp, dp, ctypes.c_double]
        L.orc_left_eigenvectors.argtypes = [dp, dp, ctypes.c_double, ctypes.c_int]
        L.orc_right_eigenvectors.argtypes = [dp, dp, ctypes.c_double, ctypes.c_int]
        L.orc_weno5_interp.argtypes = [dp, ctypes.c_double]; L.orc_weno5_interp.restype = ctypes.c_double
        L.orc_first_derivative4_1d.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        L.orc_first_derivative4_1d.restype = None
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """Oracle bound to one workload (grid, physics, scheme)."""

    def __init__(self, dims, interp="components", upwind="rusanov", Re=-1.0, Minf=1.0,
                 Pr=0.72, gamma=1.4, eps=1e-6, ghosts=3, weno_type="js"):
        from hypar_amd import decks
        self.dims = tuple(int(d) for d in dims)
        self.g = ghosts
        self.dxinv = np.ascontiguousarray(decks.dxinv_all(self.dims, ghosts))
        self.p = OrcParams()
        self.p.dim[:] = list(self.dims)
        self.p.ghosts = ghosts
        self.p.gamma = gamma
        self.p.Re = Re / Minf
        self.p.Pr = Pr
        self.p.interp = INTERP[interp]
        self.p.upwind = UPWIND[upwind]
        self.p.eps = eps
        self.p.dxinv = _dp(self.dxinv)
        self.p.weno_type = WENO_TYPES[weno_type]

    @classmethod
    def for_workload(cls, w):
        return cls(w.dims, w.interp, w.upwind, w.Re, w.Minf, w.Pr, w.gamma, w.weno_eps,
                   weno_type=w.weno_type)

    # layout helpers ------------------------------------------------------
    def ghosted(self, u):
        """(nz,ny,nx,5) interior -> reference ghosted AoS array (flat)."""
        g = self.g
        nx, ny, nz = self.dims
        G = np.zeros((nz + 2 * g, ny + 2 * g, nx + 2 * g, 5))
        G[g:g + nz, g:g + ny, g:g + nx, :] = u
        return np.ascontiguousarray(G)

    def interior(self, G):
        g = self.g
        nx, ny, nz = self.dims
        return np.ascontiguousarray(G[g:g + nz, g:g + ny, g:g + nx, :])

    # calls ---------------------------------------------------------------
    def rk4(self, u, dt, nsteps):
        return self.step(u, dt, nsteps, "44")

    # the other explicit tableaux of TimeExplicitRKInitialize.c:26-57,
    # (A row-major, b)
    TABLEAUX = {"1fe": ([0.0], [1.0]),
                "22": ([0.0, 0.0, 1.0, 0.0], [0.5, 0.5]),
                "33": ([0.0, 0.0, 0.0, 2.0 / 3.0, 0.0, 0.0, 2.0 / 3.0 - 1.0 / 4.0, 1.0 / 4.0, 0.0],
                       [1.0 / 4.0, -1.0 / 4.0, 1.0])}

    def step(self, u, dt, nsteps, scheme="44"):
        """nsteps steps of the rk-class scheme named as in solver.inp's
        time_scheme_type ("44", "ssprk3"/"tvdrk3", "1fe", "22", "33")."""
        G = self.ghosted(u)
        if scheme in self.TABLEAUX:
            A, b = (np.ascontiguousarray(x, dtype=np.float64) for x in self.TABLEAUX[scheme])
            err = lib().orc_rk_tableau(ctypes.byref(self.p), _dp(G), float(dt), int(nsteps), len(b), _dp(A), _dp(b))
        else:
            fn = {"44": lib().orc_rk4, "ssprk3": lib().orc_ssprk3, "tvdrk3": lib().orc_ssprk3}[scheme]
            err = fn(ctypes.byref(self.p), _dp(G), float(dt), int(nsteps))
        if err:
            raise RuntimeError(f"oracle {scheme} step failed")
        return self.interior(G)

    def rhs(self, u):
        G = self.ghosted(u)
        R = np.zeros_like(G)
        if lib().orc_rhs(ctypes.byref(self.p), _dp(R), _dp(G)):
            raise RuntimeError("orc_rhs failed")
        return self.interior(R)

    def hyperbolic(self, u):
        G = self.ghosted(u)
        lib().orc_periodic_bc(ctypes.byref(self.p), _dp(G))
        H = np.zeros_like(G)
        lib().orc_hyperbolic(ctypes.byref(self.p), _dp(H), _dp(G))
        return self.interior(H)

    def parabolic(self, u):
        G = self.ghosted(u)
        lib().orc_periodic_bc(ctypes.byref(self.p), _dp(G))
        P = np.zeros_like(G)
        lib().orc_parabolic(ctypes.byref(self.p), _dp(P), _dp(G))
        return self.interior(P)

    def cfl(self, u, dt):
        G = self.ghosted(u)
        return float(lib().orc_cfl(ctypes.byref(self.p), _dp(G), float(dt)))

    def volume_integral(self, u):
        G = self.ghosted(u)
        out = np.zeros(5)
        lib().orc_volume_integral(ctypes.byref(self.p), _dp(G), _dp(out))
        return out
