"""ctypes binding of libh3d.so (the C ABI in include/h3d.h and include/h3d_model.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback for the device path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libh3d.so")

# status codes and enums (include/h3d.h)
OK = 0
ERR_ARG, ERR_HIP, ERR_NAN, ERR_UNSUPPORTED, ERR_STATE = -1, -2, -3, -4, -5
ISOTHERMAL, HEAT_FLUX, CONVECTION = 0, 1, 2
SOLVER_PBICGSTAB, SOLVER_CG = 0, 1
SMOOTHER_NONE, SMOOTHER_GS = 0, 1
FP64, FP32_REFINED = 0, 1
OPT_COEF_FORM, OPT_NUM_CLASSES, OPT_LAYOUT = 0, 1, 2
LAYOUT_AUTO, LAYOUT_STANDARD = 0, 1
OPT_OPERATOR_FORM, FORM_REFERENCE, FORM_FLUX = 3, 0, 1
OPT_SWEEP_CHUNK = 4
OPT_SUM_ORDER, SUM_EXACT, SUM_SEQUENTIAL = 5, 0, 1
OPT_RESIDENT, OPT_RESIDENT_WGS = 6, 7
COEF_AUTO, COEF_DIRECT = 0, 1

FIELDS = ["theta", "b", "mask", "rho", "lam", "cp", "pcg_p", "pcg_p_", "pcg_r", "pcg_r0",
          "pcg_q", "pcg_s", "pcg_s_", "pcg_t_", "hsrc"]
FIELD_ID = {name: i for i, name in enumerate(FIELDS)}

# every exported symbol the headers declare (checked by the CPU test suite)
EXPORTS = [
    "h3d_create", "h3d_destroy", "h3d_last_error", "h3d_create_error", "h3d_set_grid", "h3d_set_bc",
    "h3d_get_hc", "h3d_upload", "h3d_download", "h3d_fill", "h3d_copy", "h3d_apply_bc", "h3d_calc_rhs",
    "h3d_pbicgstab", "h3d_cg", "h3d_get_history", "h3d_last_solve_ms", "h3d_last_solve_floor",
    "h3d_calc_ax", "h3d_calc_rk", "h3d_residual_scale",
    "h3d_rbsor_color", "h3d_precond", "h3d_fdot1", "h3d_fdot2", "h3d_bicg1", "h3d_triad", "h3d_bicg2",
    "h3d_time_kernel", "h3d_set_option", "h3d_get_option",
    "h3d_objective", "h3d_adjoint_solve", "h3d_gradients", "h3d_get_adjoint",
    "h3d_adjoint_step", "h3d_gradients_accumulate", "h3d_gradients_accumulated",
    "h3d_store_field", "h3d_load_field", "h3d_release_slots", "h3d_mem_info",
    "h3d_res_sor", "h3d_rbsor", "h3d_rbsor_solve", "h3d_sor", "h3d_fdot2_noalloc",
    "h3d_model_pitch", "h3d_model_genz", "h3d_model_fill_id", "h3d_model_properties",
    "h3d_model_heat_src", "h3d_model_mode3_bc", "h3d_model_power_map",
]

_c_int, _c_double, _c_size, _vp = ctypes.c_int, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)

_SIGS = {
    "h3d_create": (_c_int, [ctypes.POINTER(_vp), _c_int, _c_int, _c_int, _c_int]),
    "h3d_destroy": (_c_int, [_vp]),
    "h3d_last_error": (ctypes.c_char_p, [_vp]),
    "h3d_create_error": (ctypes.c_char_p, []),
    "h3d_set_grid": (_c_int, [_vp, _c_double, _c_double, _vp, _vp, _c_int]),
    "h3d_set_bc": (_c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "h3d_get_hc": (_c_int, [_vp, _vp]),
    "h3d_upload": (_c_int, [_vp, _c_int, _vp, _c_size]),
    "h3d_download": (_c_int, [_vp, _c_int, _vp, _c_size]),
    "h3d_fill": (_c_int, [_vp, _c_int, _c_double]),
    "h3d_copy": (_c_int, [_vp, _c_int, _c_int]),
    "h3d_apply_bc": (_c_int, [_vp]),
    "h3d_calc_rhs": (_c_int, [_vp, _c_double, _c_int]),
    "h3d_pbicgstab": (_c_int, [_vp, _c_double, _c_double, _c_int, _c_int, _c_int, _c_int, _c_double,
                               _ip, _ip, _dp]),
    "h3d_cg": (_c_int, [_vp, _c_double, _c_double, _c_int, _c_int, _c_int, _c_int, _c_double,
                        _ip, _ip, _dp]),
    "h3d_get_history": (_c_int, [_vp, _vp, _c_int, _ip]),
    "h3d_last_solve_ms": (_c_int, [_vp, _dp]),
    "h3d_last_solve_floor": (_c_int, [_vp, _ip]),
    "h3d_calc_ax": (_c_int, [_vp, _c_int, _c_int, _c_double, _c_int]),
    "h3d_calc_rk": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_double, _c_int, _dp]),
    "h3d_residual_scale": (_c_int, [_vp, _c_int, _c_int, _c_double, _c_int, _dp]),
    "h3d_rbsor_color": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_double, _c_int]),
    "h3d_precond": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_double, _c_int]),
    "h3d_fdot1": (_c_int, [_vp, _c_int, _dp]),
    "h3d_fdot2": (_c_int, [_vp, _c_int, _c_int, _dp]),
    "h3d_fdot2_noalloc": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _vp, _vp]),
    "h3d_bicg1": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_double, _c_double]),
    "h3d_triad": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_double]),
    "h3d_bicg2": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_double, _c_double]),
    "h3d_time_kernel": (_c_int, [_vp, _c_int, _c_int, _c_double, _c_int, _dp]),
    "h3d_objective": (_c_int, [_vp, _vp, _c_size, _c_double, _dp]),
    "h3d_adjoint_solve": (_c_int, [_vp, _c_int, _c_double, _c_double, _c_int, _c_int, _c_int, _ip, _ip, _dp]),
    "h3d_gradients": (_c_int, [_vp, _vp, _vp, _c_size]),
    "h3d_get_adjoint": (_c_int, [_vp, _vp, _vp, _c_size]),
    "h3d_adjoint_step": (_c_int, [_vp, _c_int, _c_double, _c_double, _c_int, _c_int, _c_int, _c_double, _c_int,
                                  _ip, _ip, _dp]),
    "h3d_gradients_accumulate": (_c_int, [_vp, _c_double, _c_double]),
    "h3d_gradients_accumulated": (_c_int, [_vp, _vp, _vp, _c_size, _c_int]),
    "h3d_store_field": (_c_int, [_vp, _c_int, _c_int]),
    "h3d_load_field": (_c_int, [_vp, _c_int, _c_int, _c_double]),
    "h3d_release_slots": (_c_int, [_vp]),
    "h3d_mem_info": (_c_int, [_vp, ctypes.POINTER(_c_size), ctypes.POINTER(_c_size)]),
    "h3d_res_sor": (_c_int, [_vp, _c_int, _c_int, _c_double, _c_double, _c_int, _dp]),
    "h3d_rbsor": (_c_int, [_vp, _c_int, _c_int, _c_double, _c_double, _c_int, _dp]),
    "h3d_rbsor_solve": (_c_int, [_vp, _c_double, _c_double, _c_double, _c_int, _c_int, _ip, _ip, _dp]),
    "h3d_sor": (_c_int, [_vp, _c_int, _c_int, _c_double, _c_double, _c_int, _dp]),
    "h3d_set_option": (_c_int, [_vp, _c_int, _c_int]),
    "h3d_get_option": (_c_int, [_vp, _c_int, _ip]),
    "h3d_model_pitch": (_c_int, [_c_int, _dp]),
    "h3d_model_genz": (_c_int, [_c_int, _vp, _vp, _vp]),
    "h3d_model_fill_id": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_double, _c_double, _vp, _c_int]),
    "h3d_model_properties": (_c_int, [_vp, _vp, _vp, _vp, _c_size]),
    "h3d_model_heat_src": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_double]),
    "h3d_model_mode3_bc": (_c_int, [_vp, _vp, _vp, _vp, _vp]),
    "h3d_model_power_map": (_c_int, [_vp, _vp, _c_int, _c_int, _c_int, _c_double, _c_double, _vp,
                                     ctypes.c_uint64, _c_double]),
}

_lib = None


def load():
    """Load libh3d.so once; raise RuntimeError if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libh3d.so is not built (%s); run `python -m h2_amd.build` "
                           "or __graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a):
    """Data pointer of a contiguous numpy array; 3-D arrays must use the
    reference's column-major layout (Fortran order, i fastest)."""
    if a.ndim == 3 and not a.flags["F_CONTIGUOUS"]:
        raise ValueError("3-D arrays must be Fortran-ordered (column-major, i fastest)")
    if not (a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]):
        raise ValueError("array must be contiguous")
    return ctypes.c_void_p(a.ctypes.data)


class H3DError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("libh3d error %d: %s" % (code, msg))
        self.code = code
