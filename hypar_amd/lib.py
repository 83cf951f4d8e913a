"""ctypes binding of libhypar_mi355x.so (C-ABI in include/hypar_mi355x.h).

torch is imported before the library is loaded so that the process has a
single HIP runtime (torch's bundled libamdhip64.so.7 satisfies the library's
dependency by soname).  A missing library is an error, never a fallback.
"""
from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (must precede the CDLL load, see module docstring)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhypar_mi355x.so")

HM_OK = 0
HM_ERR_ARG, HM_ERR_HIP, HM_ERR_NONFINITE, HM_ERR_UNSUPPORTED, HM_ERR_NOMEM, HM_ERR_TIMEOUT = 1, 2, 3, 4, 5, 6
INTERP = {"components": 0, "characteristic": 1}
UPWIND = {"roe": 0, "rusanov": 1}
WENO = {"js": 0, "mapped": 1, "borges": 2, "yc": 3, "optimal": 4}
LAYOUT_REFERENCE, LAYOUT_AOS_INTERIOR, LAYOUT_SOA_INTERIOR = 0, 1, 2


class HMConfig(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int * 3), ("gamma", ctypes.c_double), ("Re", ctypes.c_double),
                ("Minf", ctypes.c_double), ("Pr", ctypes.c_double), ("interp", ctypes.c_int),
                ("upwind", ctypes.c_int), ("weno_eps", ctypes.c_double),
                ("dxinv", ctypes.POINTER(ctypes.c_double)), ("weno_type", ctypes.c_int)]


class HMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hm error {code}: {msg}")
        self.code = code


_LIB = None

# (name, argtypes, restype); every symbol declared in include/*.h
_VP, _DP, _SZ = ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t
SIGNATURES = [
    ("hm_create", [ctypes.POINTER(HMConfig), ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    ("hm_destroy", [_VP], None),
    ("hm_last_error", [], ctypes.c_char_p),
    ("hm_version", [], ctypes.c_char_p),
    ("hm_set_stream", [_VP, _VP], ctypes.c_int),
    ("hm_ncell", [_VP], _SZ),
    ("hm_upload", [_VP, _DP, _DP, ctypes.c_int], ctypes.c_int),
    ("hm_download", [_VP, _DP, _DP, ctypes.c_int], ctypes.c_int),
    ("hm_rk4", [_VP, _DP, ctypes.c_double, ctypes.c_int], ctypes.c_int),
    ("hm_ssprk3", [_VP, _DP, ctypes.c_double, ctypes.c_int], ctypes.c_int),
    ("hm_rhs", [_VP, _DP, _DP], ctypes.c_int),
    ("hm_hyperbolic", [_VP, _DP, _DP], ctypes.c_int),
    ("hm_parabolic", [_VP, _DP, _DP], ctypes.c_int),
    ("hm_cfl", [_VP, _DP, ctypes.c_double, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    ("hm_step_norm", [_VP, _DP, _DP, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    ("hm_cfl_noalloc", [ctypes.POINTER(HMConfig), _DP, ctypes.c_double, ctypes.POINTER(ctypes.c_double), _VP],
     ctypes.c_int),
    ("hm_volume_integral", [_VP, _DP, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    ("hm_profile_kernels", [_VP, _DP, ctypes.c_int, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    ("hm_malloc", [ctypes.POINTER(ctypes.c_void_p), _SZ], ctypes.c_int),
    ("hm_free", [_DP], None),
    ("hm_copy_to_device", [_DP, _DP, _SZ], ctypes.c_int),
    ("hm_copy_to_host", [_DP, _DP, _SZ], ctypes.c_int),
    ("hm_copy_device", [_VP, _DP, _DP, _SZ], ctypes.c_int),
    ("hm_synchronize", [_VP], ctypes.c_int),
    ("hm_flux", [_VP, _DP, _DP, ctypes.c_int], ctypes.c_int),
    ("hm_modified_solution", [_VP, _DP, _DP], ctypes.c_int),
    ("hm_interpolate", [_VP, _DP, _DP, _DP, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    ("hm_upwind", [_VP, _DP, _DP, _DP, _DP, _DP, _DP, ctypes.c_int], ctypes.c_int),
    ("hm_flux_difference", [_VP, _DP, _DP, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    ("hm_axpy", [_VP, _DP, ctypes.c_double, _DP, _SZ], ctypes.c_int),
    ("hm_fill", [_VP, _DP, ctypes.c_double, _SZ], ctypes.c_int),
    ("hm_rhs_vjp", [_VP, _DP, _DP, _DP], ctypes.c_int),
    ("hm_rk4_vjp", [_VP, _DP, ctypes.c_double, _DP], ctypes.c_int),
    ("hm_adjoint_forward", [_VP, _DP, ctypes.c_double, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    ("hm_adjoint_reverse", [_VP, _DP], ctypes.c_int),
    ("hm_ensemble_create", [ctypes.POINTER(HMConfig), ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)],
     ctypes.c_int),
    ("hm_ensemble_destroy", [_VP], None),
    ("hm_ensemble_ncell", [_VP], _SZ),
    ("hm_ensemble_member", [_VP, ctypes.c_int, ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)], ctypes.c_int),
    ("hm_ensemble_set_stream", [_VP, _VP], ctypes.c_int),
    ("hm_ensemble_set_dt", [_VP, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    ("hm_ensemble_rk4", [_VP, _DP, ctypes.c_int], ctypes.c_int),
    ("hm_ensemble_adjoint_forward", [_VP, _DP, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    ("hm_ensemble_adjoint_reverse", [_VP, _DP], ctypes.c_int),
    ("hm_ensemble_set_tiles", [_VP, ctypes.c_int, ctypes.POINTER(ctypes.c_uint)], ctypes.c_int),
    ("hm_ensemble_get_tiles", [_VP, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint)], ctypes.c_int),
    ("hm_ensemble_tune", [_VP, _DP, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint)],
     ctypes.c_int),
    ("hm_ensemble_set_wait_limit", [_VP, ctypes.c_uint], ctypes.c_int),
    ("hm_set_tape_budget", [_VP, _SZ], ctypes.c_int),
    ("hm_tape_info", [_VP, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(_SZ)],
     ctypes.c_int),
    ("hm_ensemble_set_tape_budget", [_VP, _SZ], ctypes.c_int),
    ("hm_ensemble_tape_info", [_VP, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(_SZ)],
     ctypes.c_int),
    # include/hypar_mi355x_slots.h
    ("hm_bind", [_VP, _VP], ctypes.c_int),
    ("hm_bind_state", [_VP, _DP, ctypes.c_int], ctypes.c_int),
    ("hm_bind_time", [_VP, _VP, ctypes.c_double, ctypes.c_int], ctypes.c_int),
    ("hm_unbind", [_VP], None),
    ("hm_slot_TimeIntegrate", [_VP], ctypes.c_int),
    ("hm_slot_RHSFunction", [_DP, _DP, _VP, _VP, ctypes.c_double], ctypes.c_int),
    ("hm_slot_HyperbolicFunction", [_DP, _DP, _VP, _VP, ctypes.c_double, ctypes.c_int, _VP, _VP], ctypes.c_int),
    ("hm_slot_ParabolicFunction", [_DP, _DP, _VP, _VP, ctypes.c_double], ctypes.c_int),
    ("hm_slot_ComputeCFL", [_VP, _VP, ctypes.c_double, ctypes.c_double], ctypes.c_double),
]

# include/hypar_mi355x_slots.h: the arrays a bound object's slots receive,
# and the tableaux of hm_bind_time
SLOT_ARRAYS_DEVICE, SLOT_ARRAYS_REFERENCE_HOST = 0, 1
TIME_RK4, TIME_SSPRK3 = 0, 1

# HyPar's slot prototypes (reference include/hypar.h:220-276,
# include/timeintegration_struct.h:96-98), for holding the hm_slot_* entry
# points as plain function pointers
FFunctionProto = ctypes.CFUNCTYPE(ctypes.c_int, _DP, _DP, ctypes.c_int, _VP, ctypes.c_double)
UpwindProto = ctypes.CFUNCTYPE(ctypes.c_int, _DP, _DP, _DP, _DP, _DP, _DP, ctypes.c_int, _VP, ctypes.c_double)
TimeIntegrateProto = ctypes.CFUNCTYPE(ctypes.c_int, _VP)
RHSFunctionProto = ctypes.CFUNCTYPE(ctypes.c_int, _DP, _DP, _VP, _VP, ctypes.c_double)
HyperbolicFunctionProto = ctypes.CFUNCTYPE(ctypes.c_int, _DP, _DP, _VP, _VP, ctypes.c_double, ctypes.c_int,
                                           FFunctionProto, UpwindProto)
ParabolicFunctionProto = ctypes.CFUNCTYPE(ctypes.c_int, _DP, _DP, _VP, _VP, ctypes.c_double)
ComputeCFLProto = ctypes.CFUNCTYPE(ctypes.c_double, _VP, _VP, ctypes.c_double, ctypes.c_double)


def slot_pointer(name: str, proto):
    """The address of the library's entry point `name` as a function
    pointer of type `proto`: what a C caller stores in a HyPar slot."""
    return proto(ctypes.cast(getattr(load(), name), ctypes.c_void_p).value)


def load(path: str = LIB_PATH):
    """Load the library (once) and declare every entry point."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(path)
        for name, args, res in SIGNATURES:
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def check(code: int):
    if code != HM_OK:
        raise HMError(code, load().hm_last_error().decode())
