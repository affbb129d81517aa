"""Loads libmmx.so and declares the C signatures of include/mmx.h."""
import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmmx.so")

_vp = C.c_void_p
_fp = C.c_void_p        # device float* (passed as an integer address)
_i = C.c_int
_l = C.c_long
_f = C.c_float
_d = C.c_double
_s = C.c_char_p

# name: (restype, argtypes)
_SIGS = {
    "mmx_last_error": (_s, []),
    "mmx_version": (_s, []),
    "mmx_init": (_i, [_i]),
    "mmx_normalize": (_i, [_fp, _fp, _fp, _fp, _i, _vp]),
    "mmx_torque": (_i, [_fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _f, _i, _vp]),
    "mmx_exchange6": (_i, [_fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _f, _i, _i, _i, _i, _i, _i,
                           _f, _f, _f, _vp]),
    "mmx_uniaxial_anisotropy": (_i, [_fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _f, _fp, _f, _fp, _f,
                                     _fp, _f, _i, _vp]),
    "mmx_zhangli": (_i, [_fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _f, _f,
                         _f, _f, _f, _f, _i, _i, _i, _f, _f, _f, _i, _i, _i, _vp]),
    "mmx_vec_madd": (_i, [_fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _f, _f, _f, _i, _vp]),
    "mmx_madd": (_i, [_fp, _fp, _fp, _f, _i, _vp]),
    "mmx_madd1": (_i, [_fp, _fp, _f, _i, _vp]),
    "mmx_madd2": (_i, [_fp, _fp, _f, _fp, _f, _i, _vp]),
    "mmx_mul": (_i, [_fp, _fp, _fp, _i, _vp]),
    "mmx_reductor_create": (_i, [C.POINTER(_vp), _l]),
    "mmx_reductor_destroy": (_i, [_vp]),
    "mmx_reductor_sum": (_i, [_vp, _fp, _l, _vp, C.POINTER(_d)]),
    "mmx_reductor_dot": (_i, [_vp, _fp, _fp, _l, _vp, C.POINTER(_d)]),
    "mmx_reductor_max": (_i, [_vp, _fp, _l, _vp, C.POINTER(_f)]),
    "mmx_reductor_min": (_i, [_vp, _fp, _l, _vp, C.POINTER(_f)]),
    "mmx_reductor_maxabs": (_i, [_vp, _fp, _l, _vp, C.POINTER(_f)]),
    "mmx_reductor_maxdiff": (_i, [_vp, _fp, _fp, _l, _vp, C.POINTER(_f)]),
    "mmx_reductor_maxnorm": (_i, [_vp, _fp, _fp, _fp, _l, _vp, C.POINTER(_f)]),
    "mmx_reduce_noalloc": (_i, [_i, _fp, _fp, _fp, _l, _vp, C.POINTER(_d)]),
    "mmx_demag_plan_create": (_i, [C.POINTER(_vp), _i, _i, _i, _d, _d, _d, _i, _i, _i, _i, _i]),
    "mmx_demag_plan_destroy": (_i, [_vp]),
    "mmx_demag_update_b": (_i, [_vp, _fp, _fp, _fp, _fp, _fp, _fp, _fp, _f, _vp]),
    "mmx_demag_plan_info": (_i, [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_d)]),
    "mmx_set_fast_x_min_blocks": (_i, [_i, C.POINTER(_i)]),
    "mmx_demag_plan_kernel": (_i, [_vp, _vp]),
    "mmx_demag_plan_real_kernel": (_i, [_vp, _vp]),
    "mmx_sim_create": (_i, [C.POINTER(_vp), _i, _i, _i, _d, _d, _d, _i, _i, _i]),
    "mmx_sim_destroy": (_i, [_vp]),
    "mmx_sim_load": (_i, [_vp, _s]),
    "mmx_sim_set_kernel": (_i, [_vp, _i, _i]),
    "mmx_sim_setv": (_i, [_vp, _s, C.POINTER(_d), _i]),
    "mmx_sim_getv": (_i, [_vp, _s, C.POINTER(_d), _i]),
    "mmx_sim_setmask": (_i, [_vp, _s, _vp, _i]),
    "mmx_sim_setarray": (_i, [_vp, _s, _vp, _i]),
    "mmx_sim_getarray": (_i, [_vp, _s, _vp, _i]),
    "mmx_sim_device_array": (_i, [_vp, _s, C.POINTER(_vp)]),
    "mmx_sim_stream": (_i, [_vp, C.POINTER(_vp)]),
    "mmx_sim_update": (_i, [_vp, _s]),
    "mmx_sim_step": (_i, [_vp]),
    "mmx_sim_steps": (_i, [_vp, _i]),
    "mmx_sim_run": (_i, [_vp, _d]),
    "mmx_sim_try_step": (_i, [_vp, _d, C.POINTER(_d)]),
    "mmx_sim_time_passes": (_i, [_vp, _i, C.POINTER(_d), C.POINTER(_i)]),
    "mmx_crc64_iso": (C.c_ulonglong, [_vp, C.c_ulonglong, C.c_ulonglong]),
    "mmx_philox4x32_10": (C.c_int, [_vp, _vp, _vp]),
    "mmx_surrogate_create": (_i, [C.POINTER(_vp), _vp, C.c_ulonglong]),
    "mmx_surrogate_destroy": (_i, [_vp]),
    "mmx_surrogate_num_params": (_i, [_vp, C.POINTER(_i)]),
    "mmx_surrogate_get_params": (_i, [_vp, _vp, _i]),
    "mmx_surrogate_set_params": (_i, [_vp, _vp, _i]),
    "mmx_surrogate_get_grads": (_i, [_vp, _vp, _i]),
    "mmx_surrogate_train_step": (_i, [_vp, _fp, _f, C.POINTER(_d)]),
    "mmx_surrogate_net": (_i, [_vp, _vp, _vp, _vp, C.POINTER(_d)]),
    "mmx_surrogate_predict_b": (_i, [_vp, _fp, _fp, _i]),
    "mmx_sim_set_surrogate": (_i, [_vp, _vp, _d, _i]),
    "mmx_batch_create": (_i, [C.POINTER(_vp), C.POINTER(_vp), _i]),
    "mmx_batch_destroy": (_i, [_vp]),
    "mmx_batch_run": (_i, [_vp, _d]),
    "mmx_batch_steps": (_i, [_vp, _i]),
    "mmx_batch_set_graph": (_i, [_vp, _i]),
    "mmx_batch_set_persistent": (_i, [_vp, _i]),
    "mmx_batch_tune": (_i, [_vp, _i, C.POINTER(_i)]),
    "mmx_batch_set_lds": (_i, [_vp, C.POINTER(_i)]),
    "mmx_batch_info": (_i, [_vp, C.POINTER(_i), C.POINTER(_i)]),
    "mmx_convplan_create": (_i, [C.POINTER(_vp), _vp, _vp, _vp]),
    "mmx_convplan_destroy": (_i, [_vp]),
    "mmx_convplan_convolve": (_i, [_vp, _fp, _fp, _vp]),
    "mmx_convplan_info": (_i, [_vp, C.POINTER(_d)]),
}

# include/mmx_libmumax2.h: libmumax2's signatures (float** = address of an
# array of device pointers, hipStream_t* = address of an array of streams);
# no return value (errors abort the process)
_pp = C.c_void_p
_LM2 = {
    "mmx_setUsedGPUs": [_vp, _i],
    "mmx_normalizeAsync": [_pp] * 4 + [_vp, _i],
    "mmx_torqueAsync": [_pp] * 10 + [_f, _vp, _i],
    "mmx_exchange6Async": [_pp] * 8 + [_f, _i, _i, _i, _i, _i, _i, _f, _f, _f, _vp],
    "mmx_uniaxialAnisotropyAsync": [_pp] * 8 + [_f, _pp, _f, _pp, _f, _pp, _f, _vp, _i],
    "mmx_zhangli_async": [_pp] * 13 + [_f] * 6 + [_i] * 3 + [_f] * 3 + [_i] * 3 + [_vp],
    "mmx_maddAsync": [_pp, _pp, _pp, _f, _vp, _i],
    "mmx_madd1Async": [_pp, _pp, _f, _vp, _i],
    "mmx_madd2Async": [_pp, _pp, _f, _pp, _f, _vp, _i],
    "mmx_vecMaddAsync": [_pp] * 9 + [_f, _f, _f, _vp, _i],
    "mmx_mulAsync": [_pp, _pp, _pp, _vp, _i],
    "mmx_cmaddAsync": [_pp, _f, _f, _pp, _pp, _vp, _i],
    "mmx_kernelMulMicromag3DAsync": [_pp] * 9 + [_vp, _i],
    "mmx_copyPad3DAsync": [_pp, _i, _i, _i, _pp, _i, _i, _i, _i, _vp],
    "mmx_copyUnPad3DAsync": [_pp, _i, _i, _i, _pp, _i, _i, _i, _i, _vp],
    "mmx_zeroArrayAsync": [_pp, _i, _vp],
    "mmx_partialSumAsync": [_pp, _pp, _i, _i, _i, _vp],
    "mmx_partialSDotAsync": [_pp, _pp, _pp, _i, _i, _i, _vp],
    "mmx_partialMaxAsync": [_pp, _pp, _i, _i, _i, _vp],
    "mmx_partialMinAsync": [_pp, _pp, _i, _i, _i, _vp],
    "mmx_partialMaxAbsAsync": [_pp, _pp, _i, _i, _i, _vp],
    "mmx_partialMaxDiffAsync": [_pp, _pp, _pp, _i, _i, _i, _vp],
    "mmx_partialMaxSumAsync": [_pp, _pp, _pp, _i, _i, _i, _vp],
    "mmx_partialMaxNorm3SqAsync": [_pp] * 4 + [_i, _i, _i, _vp],
    "mmx_partialMaxNorm3SqDiffAsync": [_pp] * 7 + [_i, _i, _i, _vp],
}
for _name, _args in _LM2.items():
    _SIGS[_name] = (None, _args)

EXPORTED_SYMBOLS = tuple(sorted(_SIGS))

MMX_KERNEL_ARNE = 0
MMX_KERNEL_NEWELL = 1
MMX_KEEP_REAL_KERNEL = 0x100

_ERRNAMES = {1: "MMX_ERR_ARG", 2: "MMX_ERR_SIZE", 3: "MMX_ERR_HIP", 4: "MMX_ERR_NOMEM",
             5: "MMX_ERR_CONVERGE", 6: "MMX_ERR_NAN", 7: "MMX_ERR_NAME", 8: "MMX_ERR_STATE"}


class MMXError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (_ERRNAMES.get(code, "?"), code, msg))
        self.code = code


_lib = None


def load_library(path=None):
    """Loads libmmx.so; raises if it is not built (no fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError("libmmx.so not built at %s: run __graft_entry__.build() "
                          "(or make -C 2_amd/csrc) first" % p)
    L = C.CDLL(p)
    for name, (res, args) in _SIGS.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = L
    return L


def lib():
    return load_library()


def check(rc):
    if rc != 0:
        raise MMXError(rc, lib().mmx_last_error().decode())
    return rc
