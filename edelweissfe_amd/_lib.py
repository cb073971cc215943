"""ctypes binding of libewfe.so (the C ABI declared in include/ewfe.h).

The library is required: importing a compute entry point without it raises, there is no CPU fallback.
"""

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libewfe.so")

EWFE_C3D8, EWFE_C3D20, EWFE_C3D8TL, EWFE_C3D20TL, EWFE_C3D20R, EWFE_C3D20RTL = 0, 1, 2, 3, 4, 5
EWFE_LINEAR_ELASTIC, EWFE_VON_MISES, EWFE_NEO_HOOKE_WA, EWFE_ST_VENANT_KIRCHHOFF = 0, 1, 2, 3
EWFE_F64, EWFE_F32, EWFE_F32_MFMA, EWFE_F16, EWFE_F32_D64, EWFE_F32_MFMA_D64 = 0, 1, 2, 3, 4, 5

ELEMENT_KINDS = {"C3D8": EWFE_C3D8, "C3D20": EWFE_C3D20, "C3D8TL": EWFE_C3D8TL, "C3D20TL": EWFE_C3D20TL,
                 "C3D20R": EWFE_C3D20R, "C3D20RTL": EWFE_C3D20RTL}
MATERIAL_KINDS = {"linearelastic": EWFE_LINEAR_ELASTIC, "vonmises": EWFE_VON_MISES,
                  "neohookeanwa": EWFE_NEO_HOOKE_WA, "stvenantkirchhoff": EWFE_ST_VENANT_KIRCHHOFF}

# Every symbol include/ewfe.h declares, with (restype, argtypes).
_vp = ctypes.c_void_p
_i32, _i64, _dbl = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
_P = ctypes.POINTER
SIGNATURES = {
    "ewfe_last_error": (ctypes.c_char_p, []),
    "ewfe_version": (ctypes.c_int, []),
    "ewfe_plan_create": (ctypes.c_int, [_P(_vp), ctypes.c_int, _i64, _vp, _i64, _vp, _i64, _P(_i64)]),
    "ewfe_plan_create_chunked": (ctypes.c_int, [_P(_vp), ctypes.c_int, _i64, _vp, _i64, _vp, _i64, _P(_i64),
                                                ctypes.c_int]),
    "ewfe_plan_pipeline_info": (ctypes.c_int, [_vp, _P(_i32), _P(_i32)]),
    "ewfe_plan_set_row_gather": (ctypes.c_int, [_vp, _i32]),
    "ewfe_plan_create_host": (ctypes.c_int, [_P(_vp), ctypes.c_int, _i64, _vp, _i64, _vp, _i64, _P(_i64),
                                             ctypes.c_int]),
    "ewfe_plan_pipeline_orders": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "ewfe_plan_pattern": (ctypes.c_int, [_vp, _vp, _vp]),
    "ewfe_plan_pattern_dev": (ctypes.c_int, [_vp, _P(_vp), _P(_vp)]),
    "ewfe_plan_element_order": (ctypes.c_int, [_vp, _vp, _P(_i32)]),
    "ewfe_state_size": (ctypes.c_int, [ctypes.c_int]),
    "ewfe_plan_info": (ctypes.c_int, [_vp, _P(_i64), _P(_i32), _P(_i32), _P(_i64)]),
    "ewfe_plan_destroy": (ctypes.c_int, [_vp]),
    "ewfe_plan_set_materials": (ctypes.c_int, [_vp, _i32, _vp, _vp]),
    "ewfe_plan_create_union": (ctypes.c_int, [_P(_vp), _vp, _i32, _i64, _vp, _P(_i64)]),
    "ewfe_csr_accumulate": (ctypes.c_int, [_vp, _i32, _vp, _vp, ctypes.c_int, _vp]),
    "ewfe_eval_assemble": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                          ctypes.c_int, _vp, _vp]),
    "ewfe_adjoint_param_grad": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ewfe_body_force": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "ewfe_qp_results": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp]),
    "ewfe_surface_pressure": (ctypes.c_int, [_vp, _vp, _vp, _i64, ctypes.c_int, _dbl, _vp, _vp]),
    "ewfe_dirichlet_rows": (ctypes.c_int, [_vp, _vp, _i64, _vp, ctypes.c_int, _vp]),
    "ewfe_spmv": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _vp, _vp]),
    "ewfe_pcg": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _vp, _vp, _i64, _dbl, ctypes.c_int, ctypes.c_int, _vp,
                                _P(ctypes.c_int), _P(_dbl), _vp]),
    "ewfe_pcg_workspace_size": (_i64, [_vp]),
    "ewfe_vals_bytes": (_i64, [_vp, ctypes.c_int]),
    "ewfe_residual": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _vp, _vp, _vp]),
    "ewfe_pcg_noalloc": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _vp, _vp, _i64, _dbl, ctypes.c_int, ctypes.c_int,
                                        _P(ctypes.c_int), _P(_dbl), _vp]),
    "ewfe_mg_create": (ctypes.c_int, [_vp, ctypes.c_int, _dbl, _P(_vp), _P(ctypes.c_int)]),
    "ewfe_mg_destroy": (None, [_vp]),
    "ewfe_mg_level_info": (ctypes.c_int, [_vp, ctypes.c_int, _P(_i64), _P(_i64)]),
    "ewfe_mg_setup": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _i64, _vp]),
    "ewfe_mg_lambda": (ctypes.c_int, [_vp, _P(_dbl), ctypes.c_int]),
    "ewfe_mg_set_fine_precision": (ctypes.c_int, [_vp, ctypes.c_int]),
    "ewfe_pcg_mg": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _vp, _vp, _i64, _dbl, ctypes.c_int, ctypes.c_int,
                                   _vp, _P(ctypes.c_int), _P(_dbl), _vp]),
}


class EwfeError(RuntimeError):
    pass


_lib = None


def lib():
    """Loads libewfe.so once. Raises if it was not built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EwfeError(f"{LIB_PATH} is missing: build it with `python -m edelweissfe_amd.build` "
                            "(hipcc, gfx950); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().ewfe_last_error()
        raise EwfeError(f"libewfe error {rc}: {msg.decode() if msg else ''}")
    return rc


def ptr(t):
    """Device or host address of a torch tensor / numpy array (None -> NULL)."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return ctypes.c_void_p(t.data_ptr())
    return ctypes.c_void_p(t.ctypes.data)
