"""ctypes binding of libresaxs_amd.so (the C ABI declared in include/rsx_capi.h).

The shared library is the product: there is no CPU fallback.  Loading fails
with an explicit error when the library has not been built.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libresaxs_amd.so")

# status codes (rsx_capi.h; the first six equal resaxs::error codes)
RSX_OK = 0
RSX_INVALID_DEVICE = -1
RSX_ALG_NOT_INITIALIZED = -2
RSX_INVALID_ARG = -3
RSX_PROGRAM_NOT_LOADED = -4
RSX_ARGS_NOT_SET = -5
RSX_NO_INITIAL_CALC = -6
RSX_HIP_ERROR = -100
RSX_OUT_OF_MEMORY = -101
RSX_NOT_SUPPORTED = -102
RSX_PDB_FILE_NOT_FOUND = -1100

RSX_DOTS_DODECA_ICOSA = 0
RSX_DOTS_FIBONACCI = 1

RSX_VARIANT_WAVE = 0
RSX_VARIANT_SHARED = 1
RSX_VARIANT_FP32 = 2
RSX_VARIANT_FP32R = 3

# every symbol include/rsx_capi.h declares, with (restype, argtypes)
_c = ctypes
_vp, _fp, _dp, _ip, _up = _c.c_void_p, _c.POINTER(_c.c_float), _c.POINTER(_c.c_double), _c.POINTER(_c.c_int), \
    _c.POINTER(_c.c_uint)
_i, _f, _ll = _c.c_int, _c.c_float, _c.c_longlong
SIGNATURES = {
    "rsx_device_count": (_i, [_ip]),
    "rsx_create": (_i, [_i, _c.POINTER(_vp)]),
    "rsx_destroy": (_i, [_vp]),
    "rsx_last_error": (_c.c_char_p, [_vp]),
    "rsx_stream": (_vp, [_vp]),
    "rsx_set_q": (_i, [_vp, _fp, _i]),
    "rsx_set_factors": (_i, [_vp, _fp, _i, _i]),
    "rsx_set_atomic_factors": (_i, [_vp, _f]),
    "rsx_set_expansion_factor": (_i, [_vp, _f]),
    "rsx_set_water_weight": (_i, [_vp, _f]),
    "rsx_set_no_rsasa_recalc": (_i, [_vp, _i]),
    "rsx_set_sasa_params": (_i, [_vp, _fp, _i, _f, _i, _i]),
    "rsx_set_rebaseline": (_i, [_vp, _i]),
    "rsx_initialize": (_i, [_vp, _i]),
    "rsx_calc_curve": (_i, [_vp, _fp, _i, _fp]),
    "rsx_calc_curve_f64": (_i, [_vp, _fp, _i, _dp]),
    "rsx_recalc_curve": (_i, [_vp, _fp, _i, _i, _i, _fp]),
    "rsx_recalc_curve_f64": (_i, [_vp, _fp, _i, _i, _i, _dp]),
    "rsx_commit": (_i, [_vp]),
    "rsx_revert": (_i, [_vp]),
    "rsx_get_rsasa": (_i, [_vp, _fp, _i]),
    "rsx_state": (_i, [_vp]),
    "rsx_calc_curve_dev_async": (_i, [_vp, _vp, _i, _vp]),
    "rsx_synchronize": (_i, [_vp]),
    "rsx_calc_rsasa": (_i, [_vp, _fp, _i, _fp, _i, _f, _fp, _i, _up, _fp]),
    "rsx_calc_rsasa_noalloc": (_i, [_vp, _i, _fp, _i, _f, _vp, _i, _vp, _vp]),
    "rsx_fibonacci_dots": (_i, [_i, _f, _f, _fp]),
    "rsx_dodeca_icosa_dots": (_i, [_f, _f, _fp]),
    "rsx_sphere_dots": (_i, [_f, _f, _f, _fp, _i, _ip]),
    "rsx_device_name": (_i, [_i, _c.c_char_p, _i]),
    "rsx_atomic_form_factors": (_i, [_fp, _i, _f, _f, _fp]),
    "rsx_atomic_radii": (_i, [_fp]),
    "rsx_calc_curve_hist": (_i, [_vp, _fp, _i, _f, _dp]),
    "rsx_fit_ensemble": (_i, [_fp, _ip, _i, _fp, _fp, _fp, _i, _fp, _fp, _fp, _i, _dp, _fp]),
    "rsx_pdb_atom_type": (_i, [_c.c_char_p, _c.c_char_p, _c.c_char_p, _i]),
    "rsx_pdb_read": (_i, [_c.c_char_p, _i, _i, _fp, _ll, _c.POINTER(_ll), _ip, _i, _ip]),
    "rsx_batch_set_frames": (_i, [_vp, _fp, _ip, _i]),
    "rsx_batch_calc": (_i, [_vp, _dp, _dp]),
    "rsx_batch_vjp": (_i, [_vp, _dp, _i, _i, _fp]),
    "rsx_time_batch_vjp": (_i, [_vp, _i, _fp]),
    "rsx_batch_calc_dev_async": (_i, [_vp, _vp, _vp]),
    "rsx_set_persistent": (_i, [_vp, _i]),
    "rsx_chi2_grad": (_i, [_vp, _fp, _i, _fp, _fp, _dp, _dp, _fp, _dp]),
    "rsx_debye_vjp": (_i, [_vp, _fp, _i, _dp, _fp]),
    "rsx_debye_vjp_dev_async": (_i, [_vp, _vp, _i, _vp, _vp]),
    "rsx_chi2_grad_dev_async": (_i, [_vp, _vp, _i, _vp, _vp, _vp, _vp, _vp]),
    "rsx_autotune": (_i, [_vp, _c.c_char_p]),
    "rsx_batch_autotune": (_i, [_vp, _c.c_char_p]),
    "rsx_batch_get_tuning": (_i, [_vp, _ip, _ip, _ip, _ip]),
    "rsx_set_tuning": (_i, [_vp, _i, _i]),
    "rsx_set_variant": (_i, [_vp, _i]),
    "rsx_get_tuning": (_i, [_vp, _ip, _ip, _ip, _ip]),
    "rsx_batch_launch": (_i, [_vp, _ip, _ip]),
    "rsx_time_forward": (_i, [_vp, _i, _fp, _fp]),
    "rsx_set_adjoint_tuning": (_i, [_vp, _i, _i, _i]),
    "rsx_time_adjoint": (_i, [_vp, _i, _fp]),
    "rsx_get_adjoint_layout": (_i, [_vp, _i, _ip, _ip, _ip]),
    "rsx_time_copy": (_i, [_vp, _ll, _i, _fp]),
    "rsx_time_sasa": (_i, [_vp, _i, _fp]),
    "rsx_dev_alloc": (_i, [_vp, _c.POINTER(_vp), _ll]),
    "rsx_dev_free": (_i, [_vp, _vp]),
    "rsx_memcpy_h2d": (_i, [_vp, _vp, _vp, _ll]),
    "rsx_memcpy_d2h": (_i, [_vp, _vp, _vp, _ll]),
}

_lib = None


class RsxError(RuntimeError):
    """A non-zero status from the C ABI; .code carries the value."""

    def __init__(self, code, msg):
        super().__init__(f"rsx error {code}: {msg}")
        self.code = code


def lib():
    """Load libresaxs_amd.so (built by resaxs_amd/build.py); raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m resaxs_amd.build` "
                               "(or __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
