"""ctypes binding of libdugma_hip.so (the C ABI declared in include/dugma.h).

The library is loaded from this package directory; there is no fallback: if
it is missing or does not load, every call raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdugma_hip.so")

_d = ctypes.POINTER(ctypes.c_double)
_f = ctypes.POINTER(ctypes.c_float)
_i = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_ci = ctypes.c_int
_cd = ctypes.c_double

class BatchOpts(ctypes.Structure):
    """dugma_batch_opts of include/dugma.h."""
    _fields_ = [("max_iter", ctypes.c_int), ("anneal", ctypes.c_int),
                ("anneal_decay", ctypes.c_double), ("tol_rot", ctypes.c_double),
                ("tol_trans", ctypes.c_double), ("use_graph", ctypes.c_int),
                ("one_trial_per_launch", ctypes.c_int), ("check_every", ctypes.c_int),
                ("grid", ctypes.c_int), ("sigma_update", ctypes.c_int), ("relax", ctypes.c_double),
                ("coarse_iters", ctypes.c_int), ("refine_iters", ctypes.c_int),
                ("refine_coarse_iters", ctypes.c_int), ("refine_mid_iters", ctypes.c_int)]


class RegisterOpts(ctypes.Structure):
    """dugma_register_opts of include/dugma.h."""
    _fields_ = [("max_iter", ctypes.c_int), ("acc_rot", ctypes.c_double),
                ("acc_trans", ctypes.c_double), ("acc_sig", ctypes.c_double),
                ("verbose", ctypes.c_int)]


class RegisterEnergyOpts(ctypes.Structure):
    """dugma_register_energy_opts of include/dugma.h."""
    _fields_ = [("max_iter", ctypes.c_int), ("tol_rot", ctypes.c_double),
                ("tol_trans", ctypes.c_double), ("cov_mode", ctypes.c_int)]


_pp = ctypes.POINTER(ctypes.POINTER(ctypes.c_double))

# name: (restype, argtypes) -- one entry per declaration of include/dugma.h
SIGNATURES = {
    "dugma_init": (_ci, [_d, _d, _d, _d, _ci, _ci, _ci, ctypes.POINTER(_vp), _d]),
    "dugma_update_sig": (_ci, [_vp, _d, _d, _cd, _cd, _cd, _d]),
    "dugma_compi": (_ci, [_vp, _d, _d, _cd, _cd, _cd, _cd, _d]),
    "dugma_fr_m": (_ci, [_vp, _d]),
    "dugma_poly_eval": (_cd, [_ci, _d, _d, _d]),
    "dugma_poly_minimize": (_ci, [_ci, _d, _d, _d]),
    "dugma_register": (_ci, [_d, _d, _d, _d, _ci, _ci, _ci, _d, _d, ctypes.POINTER(RegisterOpts), _d, _d, _i, _d]),
    "dugma_energy": (_ci, [_vp, _d, _d, _ci, _d]),
    "dugma_energy_grad": (_ci, [_vp, _d, _d, _ci, _d, _d]),
    "dugma_energy_grad_noalloc": (_ci, [_vp, _d, _d, _ci, _d, _d]),
    "dugma_register_energy": (_ci, [_vp, _d, _d, ctypes.POINTER(RegisterEnergyOpts), _d, _d, _i, _d]),
    "dugma_energy_grad_points": (_ci, [_vp, _d, _d, _ci, _d, _d, _d, _d, _d, _d, _d]),
    "dugma_energy_grad_fd": (_ci, [_vp, _d, _d, _ci, _cd, _d, _d]),
    "dugma_energy_grad_hess": (_ci, [_vp, _d, _d, _ci, _d, _d, _d]),
    "dugma_knn": (_ci, [_vp, _ci, _ci, _d, _d, _i, _f]),
    "dugma_update_sigma_knn": (_ci, [_vp, _ci, _ci, _f, _f, _i]),
    "dugma_update_cov_knn": (_ci, [_vp, _ci, _ci, ctypes.c_double, ctypes.c_double, _f, _f, _i]),
    "dugma_autotune": (_ci, [_vp, _ci]),
    "dugma_ctx_plan": (_ci, [_vp, _ci, _i, _i]),
    "dugma_batch_create": (_ci, [_ci, _ci, _i, _i, _pp, _pp, _pp, _pp, ctypes.POINTER(_vp)]),
    "dugma_batch_create_iso": (_ci, [_ci, _i, _i, _pp, _pp, _pp, _pp, ctypes.POINTER(_vp)]),
    "dugma_batch_set_covariances": (_ci, [_vp, _pp, _pp]),
    "dugma_batch_register": (_ci, [_vp, _d, _d, ctypes.POINTER(BatchOpts), _d, _d, _i, _i, _d]),
    "dugma_batch_info": (_ci, [_vp, _i, _i, ctypes.POINTER(ctypes.c_longlong)]),
    "dugma_batch_free": (_ci, [_vp]),
    "dugma_batch_autotune": (_ci, [_vp, _d, _d, _ci, _i, _i, _i, _d, _d]),
    "dugma_batch_get_shape": (_ci, [_vp, _ci, _i, _i, _i]),
    "dugma_batch_set_level_shape": (_ci, [_vp, _ci, _ci, _i, _ci]),
    "dugma_batch_set_shape": (_ci, [_vp, _ci, _i, _ci]),
    "dugma_ctx_handles": (_ci, [_vp, ctypes.POINTER(_vp)]),
    "dugma_read_handle": (_ci, [_vp, _ci, _f]),
    "dugma_ctx_dims": (_ci, [_vp, _i, _i, _i]),
    "dugma_ctx_energy_info": (_ci, [_vp, _i, _i, _i]),
    "dugma_ctx_stream": (_vp, [_vp]),
    "dugma_ctx_set_timing": (_ci, [_vp, _ci]),
    "dugma_ctx_pair_time": (_ci, [_vp, _d, ctypes.POINTER(ctypes.c_longlong)]),
    "dugma_last_error": (ctypes.c_char_p, []),
    "dugma_version": (ctypes.c_char_p, []),
}

_LIB = None


class DugmaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("dugma error %d: %s" % (code, msg))
        self.code = code


def load():
    """Load libdugma_hip.so once and bind every C-ABI symbol."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libdugma_hip.so is not built (run __graft_entry__.build() "
                           "or python dugma_amd/build.py)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(code):
    if code != 0:
        msg = load().dugma_last_error()
        raise DugmaError(code, msg.decode() if msg else "")
    return code
