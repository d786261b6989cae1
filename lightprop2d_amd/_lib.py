"""ctypes binding of liblp2d.so (include/lp2d.h).

The library is built in-tree by ``lightprop2d_amd._build`` and loaded from the package
directory.  There is no fallback: if the library is missing or cannot be loaded,
every operation raises ``Lp2dError``.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "liblp2d.so")

LP2D_OK = 0
LP2D_E_ARG = -1
LP2D_E_SIZE = -2
LP2D_E_HIP = -3
LP2D_E_DTYPE = -4
LP2D_E_INIT = -5

LP2D_C64 = 0
LP2D_C128 = 1

LP2D_MASK_U8 = 1
LP2D_MASK_F32 = 2
LP2D_MASK_F64 = 3
LP2D_MASK_C64 = 4
LP2D_MASK_C128 = 5

EVANESCENT = {"reference": 0, "pass": 0, "zero": 1, "decay": 2}

LP2D_ADJ_NO_GRAD_IN = 1

# operations of lp2d_plan_workspace_size
LP2D_OP_FFT2, LP2D_OP_PROPAGATE, LP2D_OP_ADJOINT, LP2D_OP_STACK, LP2D_OP_STACK_ADJOINT = range(5)

# every symbol include/lp2d.h declares
EXPORTED = [
    "lp2d_init", "lp2d_last_error", "lp2d_version", "lp2d_max_n", "lp2d_kz_grid",
    "lp2d_fft2", "lp2d_spatial_apply", "lp2d_spectral_apply", "lp2d_intensity",
    "lp2d_propagate", "lp2d_propagate_adjoint", "lp2d_propagate_adjoint_ws",
    "lp2d_propagate_adjoint_noalloc", "lp2d_propagate_stack",
    "lp2d_propagate_stack_adjoint", "lp2d_propagate_stack_adjoint_ws_size",
    "lp2d_plan_create", "lp2d_plan_destroy", "lp2d_plan_workspace_size", "lp2d_plan_fft2",
    "lp2d_plan_propagate", "lp2d_plan_propagate_adjoint", "lp2d_plan_propagate_stack",
    "lp2d_plan_propagate_stack_adjoint", "lp2d_phase_grad",
    "lp2d_sum_planes", "lp2d_batch_create",
    "lp2d_batch_run", "lp2d_batch_destroy", "lp2d_variant_count",
    "lp2d_set_variant", "lp2d_get_variant", "lp2d_variant_name", "lp2d_trace_enable",
    "lp2d_trace_collect", "lp2d_moments", "lp2d_set_chunk", "lp2d_get_chunk",
    "lp2d_profile", "lp2d_expand", "lp2d_coarse", "lp2d_interp_linear", "lp2d_modes_project",
    "lp2d_modes_project_ws_size", "lp2d_modes_project_ws",
    "lp2d_modes_construct",
]

# generator kinds of lp2d_profile
PROF = {"plane": 0, "gaussian": 1, "round_hole": 2, "rect_hole": 3, "slits": 4,
        "random": 5, "random_hole": 6, "random_phase": 7}
LP2D_MASK_U8, LP2D_MASK_F64, LP2D_MASK_C128 = 1, 3, 5


class Lp2dError(RuntimeError):
    """A liblp2d call failed (negative status) or the library is unavailable."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


class lp2d_mask(C.Structure):
    _fields_ = [("ptr", C.c_void_p), ("dtype", C.c_int32), ("reserved", C.c_int32),
                ("stride_b", C.c_int64), ("stride_y", C.c_int64), ("stride_x", C.c_int64)]


class lp2d_optics(C.Structure):
    _fields_ = [("area_size", C.c_double), ("wl", C.c_double), ("z", C.c_double),
                ("aperture_radius", C.c_double), ("lens_f", C.c_double),
                ("evanescent", C.c_int32), ("reserved", C.c_int32)]


class lp2d_field_desc(C.Structure):
    _fields_ = [("n", C.c_int32), ("reserved", C.c_int32), ("src", C.c_void_p),
                ("dst", C.c_void_p), ("area_size", C.c_double), ("wl", C.c_double),
                ("z", C.c_double), ("aperture_radius", C.c_double)]


class lp2d_profile_params(C.Structure):
    _fields_ = [("a0", C.c_double), ("rho0", C.c_double), ("r", C.c_double),
                ("thresh", C.c_double), ("x0", C.c_double), ("y0", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double), ("slits", C.c_double),
                ("binning", C.c_int32), ("reserved", C.c_int32)]


class lp2d_interp_desc(C.Structure):
    _fields_ = [("win_off", C.c_int32), ("win_len", C.c_int32), ("src_i0", C.c_int32),
                ("dst_i0", C.c_int32), ("dst_sign", C.c_int32), ("grid_i0", C.c_int32),
                ("src_dx", C.c_double), ("dst_dx", C.c_double), ("dst_scale", C.c_double),
                ("scale", C.c_double), ("phase_coef", C.c_double), ("grid_dl", C.c_double)]


class lp2d_trace_stat(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("launches", C.c_int32), ("reserved", C.c_int32),
                ("total_ms", C.c_double)]


_lib = None
_load_error = None

_P = C.c_void_p
_I = C.c_int
_I64 = C.c_int64

_SIGS = {
    "lp2d_init": (_I, [_I]),
    "lp2d_last_error": (C.c_char_p, []),
    "lp2d_version": (_I, []),
    "lp2d_max_n": (_I, [_I]),
    "lp2d_kz_grid": (_I, [_I, C.POINTER(lp2d_optics), _P, _P]),
    "lp2d_fft2": (_I, [_I, _I, _I, _P, _P, _I, _P]),
    "lp2d_spatial_apply": (_I, [_I, _I, _I, _P, C.POINTER(lp2d_mask), _I,
                                C.POINTER(lp2d_optics), _P, _I64, _P, _P]),
    "lp2d_spectral_apply": (_I, [_I, _I, _I, _P, C.POINTER(lp2d_mask), _I,
                                 C.POINTER(lp2d_optics), _P, _P]),
    "lp2d_intensity": (_I, [_I64, _I, _P, _P, _P]),
    "lp2d_propagate": (_I, [_I, _I, _I, _P, _P, _P, _I64, C.POINTER(lp2d_optics), _P, _P]),
    "lp2d_propagate_adjoint": (_I, [_I, _I, _I, _P, _P, _P, _I64, C.POINTER(lp2d_optics),
                                    _P, _P, _I, _P]),
    "lp2d_propagate_adjoint_ws": (_I, [_I, _I, _I, _P, _P, _P, _I64, C.POINTER(lp2d_optics),
                                       _P, _P, _I, _P, _P]),
    "lp2d_propagate_adjoint_noalloc": (_I, [_I, _I, _I, _P, _P, _P, _I64,
                                            C.POINTER(lp2d_optics), _P, _P, _I, _P]),
    "lp2d_propagate_stack": (_I, [_I, _I, _I, _P, _P, _P, _I, C.POINTER(lp2d_optics),
                                  C.POINTER(_P), _P]),
    "lp2d_propagate_stack_adjoint": (_I, [_I, _I, _I, _P, C.POINTER(_P), _P, _I,
                                          C.POINTER(lp2d_optics), _P, _P, _I, _P, _I64, _P]),
    "lp2d_propagate_stack_adjoint_ws_size": (_I64, [_I, _I, _I]),
    "lp2d_plan_create": (_I, [_I, _I, _I, C.POINTER(_P)]),
    "lp2d_plan_destroy": (_I, [_P]),
    "lp2d_plan_workspace_size": (_I64, [_P, _I]),
    "lp2d_plan_fft2": (_I, [_P, _P, _P, _I, _P]),
    "lp2d_plan_propagate": (_I, [_P, _P, _P, _P, _I64, C.POINTER(lp2d_optics), _P, _P]),
    "lp2d_plan_propagate_adjoint": (_I, [_P, _P, _P, _P, _I64, C.POINTER(lp2d_optics), _P, _P,
                                         _I, _P, _I64, _P]),
    "lp2d_plan_propagate_stack": (_I, [_P, _P, _P, _P, _I, C.POINTER(lp2d_optics),
                                       C.POINTER(_P), _P]),
    "lp2d_plan_propagate_stack_adjoint": (_I, [_P, _P, C.POINTER(_P), _P, _I,
                                               C.POINTER(lp2d_optics), _P, _P, _I, _P, _I64,
                                               _P]),
    "lp2d_sum_planes": (_I, [_I64, _I, _P, _P, _P]),
    "lp2d_phase_grad": (_I, [_I, _I, _I, _P, _P, _P, _I64, _P, _I, _P]),
    "lp2d_batch_create": (_I, [C.POINTER(lp2d_field_desc), _I, _I, C.POINTER(_P)]),
    "lp2d_batch_run": (_I, [_P, _P]),
    "lp2d_batch_destroy": (_I, [_P]),
    "lp2d_variant_count": (_I, [_I, _I, _I]),
    "lp2d_set_variant": (_I, [_I, _I, _I, _I]),
    "lp2d_get_variant": (_I, [_I, _I, _I]),
    "lp2d_variant_name": (_I, [_I, _I, _I, _I, C.c_char_p, _I]),
    "lp2d_trace_enable": (_I, [_I]),
    "lp2d_moments": (_I, [_I, _I, _I, _P, C.c_double, _P, _P]),
    "lp2d_set_chunk": (_I, [_I, _I, _I]),
    "lp2d_get_chunk": (_I, [_I, _I]),
    "lp2d_trace_collect": (_I, [C.POINTER(lp2d_trace_stat), _I, C.POINTER(_I)]),
    "lp2d_profile": (_I, [_I, _I, _I, _P, _P, C.POINTER(lp2d_profile_params), _P, _P, _I, _P]),
    "lp2d_expand": (_I, [_I, _I, _I, _I, _P, _P, _P]),
    "lp2d_coarse": (_I, [_I, _I, _I, _I, _P, _P, _P]),
    "lp2d_interp_linear": (_I, [_I, _I, _I, _I, _P, _P, C.POINTER(lp2d_interp_desc), _P]),
    "lp2d_modes_project": (_I, [_I, _I, _I, _I, _I, _P, _P, _P, _P]),
    "lp2d_modes_project_ws_size": (C.c_longlong, [_I, _I, _I, _I]),
    "lp2d_modes_project_ws": (_I, [_I, _I, _I, _I, _I, _P, _P, _P, _P, C.c_longlong, _P]),
    "lp2d_modes_construct": (_I, [_I, _I, _I, _I, _I, _P, _P, _P, _P]),
}


def load(path: str | None = None):
    """Load liblp2d.so (once).  Raises Lp2dError when it is missing."""
    global _lib, _load_error
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        _load_error = (f"{p} not found: build it with `python -m lightprop2d_amd._build` "
                       "(the HIP library is required; there is no CPU fallback)")
        raise Lp2dError(_load_error)
    try:
        lib = C.CDLL(p)
    except OSError as e:  # pragma: no cover - depends on the ROCm runtime
        _load_error = f"cannot load {p}: {e}"
        raise Lp2dError(_load_error) from e
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str = "") -> None:
    if rc != LP2D_OK:
        msg = _lib.lp2d_last_error().decode() if _lib is not None else ""
        raise Lp2dError(f"{what}: liblp2d status {rc}: {msg}".strip(": "), rc)


def lib():
    return load()
