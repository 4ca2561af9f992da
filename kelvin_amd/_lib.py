"""ctypes binding of libftcc.so (the C ABI declared in include/ftcc.h).

The library is loaded from this package directory only.  There is no
fallback: if it is missing, or no GPU is present, the product path raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libftcc.so")

PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2}

_c_double_p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64

# name -> (restype, argtypes)
SIGNATURES = {
    "ftcc_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   _vp]),
    "ftcc_destroy": (ctypes.c_int, [_vp]),
    "ftcc_last_error": (ctypes.c_char_p, [_vp]),
    "ftcc_set_precision": (ctypes.c_int, [_vp, ctypes.c_int]),
    "ftcc_set_system": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp,
                                       ctypes.c_double, ctypes.c_double,
                                       ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int),
                                       _c_double_p]),
    "ftcc_set_blocks": (ctypes.c_int, [_vp, ctypes.POINTER(_vp),
                                       ctypes.POINTER(_vp)]),
    "ftcc_set_denominators": (ctypes.c_int, [_vp, _vp, _vp]),
    "ftcc_get_block": (ctypes.c_int, [_vp, ctypes.c_int, _vp]),
    "ftcc_set_grid": (ctypes.c_int, [_vp, ctypes.c_int, _c_double_p,
                                     _c_double_p]),
    "ftcc_driving_terms": (ctypes.c_int, [_vp, _vp, _vp]),
    "ftcc_residual": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp,
                                     ctypes.c_uint64]),
    "ftcc_integrate": (ctypes.c_int, [_vp, _vp, _vp]),
    "ftcc_energy": (ctypes.c_int, [_vp, _vp, _vp, _c_double_p,
                                   ctypes.c_double, ctypes.c_int,
                                   _c_double_p]),
    "ftcc_energy_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _vp,
                                           ctypes.c_double, ctypes.c_int,
                                           _vp]),
    "ftcc_sumsq": (ctypes.c_int, [_vp, _vp, _i64, _c_double_p]),
    "ftcc_step": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _c_double_p,
                                 ctypes.c_double, ctypes.c_double,
                                 ctypes.c_uint64, _c_double_p]),
    "ftcc_point_begin": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp,
                                        _vp]),
    "ftcc_point_step": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp,
                                       ctypes.c_double, _c_double_p]),
    "ftcc_td_rhs": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_int]),
    "ftcc_set_chunk": (ctypes.c_int, [_vp, ctypes.c_int]),
    "ftcc_get_chunk": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "ftcc_set_trusted": (ctypes.c_int, [_vp, ctypes.c_int]),
    "ftcc_set_correction": (ctypes.c_int, [_vp, _vp, _vp]),
    "ftcc_adjoint_begin": (ctypes.c_int, [_vp, _vp, _vp]),
    "ftcc_adjoint_cached": (ctypes.c_int, [_vp,
                                           ctypes.POINTER(ctypes.c_int)]),
    "ftcc_adjoint_apply": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "ftcc_adjoint_end": (ctypes.c_int, [_vp]),
    "ftcc_adjoint_precision": (ctypes.c_int, [_vp, ctypes.c_int]),
    "ftcc_density1": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _c_double_p,
                                     ctypes.c_double, _vp, _vp, _vp, _vp]),
    "ftcc_integrate_adjoint": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "ftcc_lambda_update": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp,
                                          _c_double_p, ctypes.c_double,
                                          ctypes.c_double, _c_double_p]),
    "ftcc_grid_derivative": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp,
                                            _c_double_p, ctypes.c_double,
                                            _c_double_p]),
    "ftcc_dot": (ctypes.c_int, [_vp, _vp, _vp, _i64, _c_double_p]),
    "ftcc_lambda_guess": (ctypes.c_int, [_vp, _vp, _c_double_p,
                                         ctypes.c_double, _vp, _vp]),
    "ftcc_lambda_convert": (ctypes.c_int, [_vp, ctypes.c_int, _c_double_p,
                                           ctypes.c_double, _vp, _vp, _vp,
                                           _vp]),
    "ftcc_set_system_response": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp,
                                                _vp, _vp, _vp,
                                                ctypes.c_double,
                                                ctypes.c_double,
                                                ctypes.POINTER(ctypes.c_int),
                                                ctypes.POINTER(ctypes.c_int)]),
    "ftcc_defect": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp, _vp,
                                   _vp, _vp, ctypes.c_uint64]),
    "ftcc_set_weights": (ctypes.c_int, [_vp, _c_double_p]),
    "ftcc_step_graph": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp,
                                       ctypes.c_double, ctypes.c_double,
                                       ctypes.c_uint64, _c_double_p]),
    "ftcc_td_step": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_int, _vp, _vp]),
    "ftcc_td_implicit_step": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_int, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_int,
                                             _vp, _vp,
                                             ctypes.POINTER(ctypes.c_int)]),
    "ftcc_ueg_eri": (ctypes.c_int, [_vp, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_int), _c_double_p,
                                    ctypes.c_double, _vp]),
    "ftcc_mean_field": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp,
                                       _c_double_p]),
    "ftcc_set_spin": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(ctypes.c_int)]),
    "ftcc_get_spin": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "ftcc_workspace_bytes": (ctypes.c_int, [_vp, ctypes.POINTER(_i64)]),
    "ftcc_profile_enable": (ctypes.c_int, [_vp, ctypes.c_int]),
    "ftcc_profile_read": (ctypes.c_int, [_vp, ctypes.POINTER(_i64),
                                         _c_double_p, _c_double_p]),
    "ftcc_set_gemm_plan": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "ftcc_profile_class": (ctypes.c_int, [_vp, ctypes.c_int,
                                          ctypes.POINTER(_i64), _c_double_p,
                                          _c_double_p, _c_double_p]),
    "ftcc_gemm": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                 _vp, _i64, _i64, _i64,
                                 _vp, _i64, _i64, _i64,
                                 ctypes.c_double,
                                 _vp, _i64, _i64, _i64, ctypes.c_int]),
}

_lib = None


class FTCCError(RuntimeError):
    """Raised when a libftcc call returns a nonzero status."""


def load():
    """Load libftcc.so from this package (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FTCCError("libftcc.so is not built; run "
                        "`python -m kelvin_amd.build` (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, ctx, what):
    if rc != 0:
        msg = load().ftcc_last_error(ctx)
        msg = msg.decode() if msg else ""
        raise FTCCError("%s failed (code %d): %s" % (what, rc, msg))
