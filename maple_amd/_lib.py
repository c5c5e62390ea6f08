"""ctypes binding of libmaple_hip.so (the C ABI declared in include/maple_hip.h).

The library is built in-tree (``python -m maple_amd.build``) at
``maple_amd/lib/libmaple_hip.so``.  Loading fails loudly when it is missing:
there is no CPU fallback in the product path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmaple_hip.so")

MAPLE_OK = 0
MAPLE_EINVAL = -1
MAPLE_EDEVICE = -2
MAPLE_ENOMEM = -3
MAPLE_ENOCONV = -4
MAPLE_ENONFINITE = -5
MAPLE_EBREAKDOWN = -6
MAPLE_ELASTIC = 1
MAPLE_THERMAL = 2
MAPLE_PREC_FP64 = 0
MAPLE_PREC_MIXED = 1
MAPLE_SMOOTH_JACOBI = 0
MAPLE_SMOOTH_L1_JACOBI = 1
MAPLE_KERNEL_FINE_OPERATOR = 0
MAPLE_KERNEL_COARSE_OPERATOR = 1
MAPLE_KERNEL_THERMAL_OPERATOR = 2

_P = C.c_void_p
_D = C.c_double
_I = C.c_int
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)

# name -> (restype, argtypes); the complete exported surface of maple_hip.h
SIGNATURES = {
    "maple_last_error": (C.c_char_p, []),
    "maple_version": (C.c_char_p, []),
    "maple_device_count": (_I, []),
    "maple_hex8_matrices": (_I, [_I, _D, _D, _D, _DP, _DP, _DP, _DP, _DP, _DP]),
    "maple_ctx_create": (_I, [_I, _I, _IP, _I, _I, C.POINTER(_P)]),
    "maple_ctx_destroy": (_I, [_P]),
    "maple_set_element_matrices": (_I, [_P, _I, _DP, _DP, _DP, _DP, _DP, _DP]),
    "maple_set_solver": (_I, [_P, _D, _I, _I]),
    "maple_set_smoother": (_I, [_P, _I]),
    "maple_set_sweeps": (_I, [_P, _I, _I]),
    "maple_set_islands": (_I, [_P, _I, _D]),
    "maple_num_islands": (_I, [_P]),
    "maple_set_cg": (_I, [_P, _I]),
    "maple_set_mixed_fallback": (_I, [_P, _I]),
    "maple_set_concurrency": (_I, [_P, _I]),
    "maple_set_graphs": (_I, [_P, _I]),
    "maple_set_chebyshev": (_I, [_P, C.c_double]),
    "maple_graph_counts": (_I, [_P, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "maple_set_coarse_refresh": (_I, [_P, _I]),
    "maple_set_start_history": (_I, [_P, _I]),
    "maple_set_variant": (_I, [_P, _I, _I]),
    "maple_homogenize": (_I, [_P, _P, _I, _D, _I, _P, _P, _IP, _P]),
    "maple_sensitivity": (_I, [_P, _P, _P, _P, _P]),
    "maple_filter_box": (_I, [_P, _P, _P, _I, _P]),
    "maple_symmetrize": (_I, [_P, _I, _I, _P, _P, _P]),
    "maple_oc_update": (_I, [_P, _P, _P, _DP, _DP, _DP, _I, _P, _P]),
    "maple_oc_update_noalloc": (_I, [_I, _I, _IP, _P, _P, _DP, _DP, _DP, _I, _P, _P]),
    "maple_penalty_overhang": (_I, [_P, _P, _D, _I, _P, _P, _P]),
    "maple_set_cavity": (_I, [_P, _P]),
    "maple_penalty_cavity": (_I, [_P, _P, _P, _P, _P]),
    "maple_rhs": (_I, [_P, _I, _P, _P, _P]),
    "maple_apply_operator": (_I, [_P, _I, _P, _I, _P, _P, _P]),
    "maple_get_fields": (_I, [_P, _I, _P, _P]),
    "maple_set_fields": (_I, [_P, _I, _P, _P]),
    "maple_set_start": (_I, [_P, _I, _P, _P]),
    "maple_start_sources": (_I, [_P, _I, _IP]),
    "maple_prolong": (_I, [_P, _I, _I, _P, _P, _P]),
    "maple_num_elements": (C.c_longlong, [_P]),
    "maple_num_nodes": (C.c_longlong, [_P, _I]),
    "maple_timing_enable": (_I, [_P, _I]),
    "maple_timing_read": (_I, [_P, _I, C.POINTER(C.c_longlong), _DP, C.POINTER(C.c_longlong)]),
    "maple_mg_setup": (_I, [_P, _I, _P, _I, _P]),
    "maple_mg_stencil": (_I, [_P, _I, _I, _I, _P, _P]),
    "maple_mg_dinv": (_I, [_P, _I, _I, _I, _P, _P]),
    "maple_mg_apply": (_I, [_P, _I, _I, _I, _I, _P, _P, _P, _P]),
    "maple_mg_restrict": (_I, [_P, _I, _I, _P, _P, _P]),
    "maple_mg_vcycle": (_I, [_P, _I, _I, _P, _P, _P]),
    "maple_mg_coarsest_inverse": (_I, [_P, _I, _I, _P, _P]),
    "maple_mg_coarsest_order": (_I, [_P, _I, _I]),
    "maple_unet_num_params": (C.c_longlong, []),
    "maple_unet_create": (_I, [_I, _I, _I, C.POINTER(_P)]),
    "maple_unet_destroy": (_I, [_P]),
    "maple_unet_set_params": (_I, [_P, C.POINTER(C.c_float)]),
    "maple_unet_get_params": (_I, [_P, C.POINTER(C.c_float)]),
    "maple_unet_set_loss": (_I, [_P, _I]),
    "maple_unet_forward": (_I, [_P, _P, _I, _P, _P]),
    "maple_unet_loss_grad": (_I, [_P, _P, _P, _I, _DP, _P, _P]),
    "maple_unet_train_step": (_I, [_P, _P, _P, _I, C.c_float, _DP, _P]),
}

SEED_FACES = {"all": 0, "x": 1, "y": 2, "z": 3, "xy": 4}


class CavityOpts(C.Structure):
    """maple_cavity_opts of maple_hip.h."""
    _fields_ = [("void_thresh", C.c_double), ("beta", C.c_double), ("min_channel_rad", C.c_int), ("rad", C.c_int),
                ("iters", C.c_int), ("periodic", C.c_int), ("seed_faces", C.c_int), ("tiles", C.c_int * 3),
                ("focus_center", C.c_int)]


_lib = None


class MapleError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"maple_hip error {code}: {msg}")
        self.code = code


class NotConverged(MapleError):
    pass


class Breakdown(NotConverged):
    """CG met p^T K p <= 0 on a live load case (results of the fields reached are written)."""


def load(path: str = LIB_PATH):
    """Load the library (once) and declare every signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -m maple_amd.build` (hipcc, gfx950)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int):
    if rc == MAPLE_OK:
        return
    msg = load().maple_last_error().decode(errors="replace")
    if rc == MAPLE_EINVAL:
        raise ValueError(f"maple_hip: {msg}")
    if rc == MAPLE_ENOCONV:
        raise NotConverged(rc, msg)
    if rc == MAPLE_EBREAKDOWN:
        raise Breakdown(rc, msg)
    raise MapleError(rc, msg)


def dptr(arr):
    """ctypes double* of a C-contiguous float64 numpy array."""
    return arr.ctypes.data_as(_DP)
