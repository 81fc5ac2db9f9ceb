"""ctypes binding of libsus2amd.so (the C ABI in include/sus2amd.h).

The library is built in-tree by superlinear_amd/build.py.  There is no CPU
fallback: if the shared object is missing or fails to load, every entry point
raises.
"""
import ctypes
import threading
from pathlib import Path

_HERE = Path(__file__).resolve().parent
_PATH = _HERE / "libsus2amd.so"
_LOCK = threading.Lock()
_LIB = None

SUS2_OK = 0
SUS2_FP64 = 0
SUS2_FP32 = 1
SUS2_MIXED = 2
STATUS = {1: "invalid", 2: "io", 3: "unsupported", 4: "device", 5: "range"}

_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)



class Sus2Objective(ctypes.Structure):
    """sus2_objective: the objective options of a model handle."""
    _fields_ = [("rel_forces", ctypes.c_double), ("norm_by_forces", ctypes.c_int), ("force_loss", ctypes.c_int),
                ("log_cosh_scale", ctypes.c_double), ("std_weight", ctypes.c_double),
                ("stdd_weight", ctypes.c_double), ("constr_weight", ctypes.c_double)]


# name -> (restype, argtypes)
SIGNATURES = {
    "sus2_objective_default": (None, [ctypes.POINTER(Sus2Objective)]),
    "sus2_objective_get": (_i, [_vp, ctypes.POINTER(Sus2Objective)]),
    "sus2_objective_set": (_i, [_vp, ctypes.POINTER(Sus2Objective)]),
    "sus2_abi_version": (_i, []),
    "sus2_last_error": (ctypes.c_char_p, []),
    "sus2_model_load": (_i, [ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "sus2_model_save": (_i, [_vp, ctypes.c_char_p]),
    "sus2_model_free": (None, [_vp]),
    "sus2_model_dims": (_i, [_vp, _ip]),
    "sus2_model_cutoff": (_d, [_vp]),
    "sus2_coeff_count": (_i, [_vp]),
    "sus2_coeff_get": (_i, [_vp, _dp]),
    "sus2_coeff_set": (_i, [_vp, _dp]),
    "sus2_coeff_offsets": (_i, [_vp, _ip]),
    "sus2_shape_build": (_i, [_i, _i, _i, _i, ctypes.c_char_p, _i]),
    "sus2_efs": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sus2_efs_host": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sus2_efs_noalloc": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sus2_graph_efs_create": (_i, [_vp, _i, _i, _i] + [_vp] * 10 + [ctypes.POINTER(_vp)]),
    "sus2_graph_loss_grad_create": (_i, [_vp, _i, _i, _i] + [_vp] * 8 + [_i, _d, _d, _d] + [_vp] * 6
                                    + [ctypes.POINTER(_vp)]),
    "sus2_graph_launch": (_i, [_vp, _vp]),
    "sus2_graph_overflow": (_i, [_vp, _vp, ctypes.POINTER(_i)]),
    "sus2_graph_free": (None, [_vp]),
    "sus2_train_bfgs": (_i, [_vp, _i, _i, _i] + [_vp] * 8 + [_i, _d, _d, _d, _i, _d, _dp, ctypes.POINTER(_i), _vp]),
    "sus2_train_linear": (_i, [_vp, _i, _i, _i] + [_vp] * 8 + [_i, _d, _d, _d, _d, _dp, _vp]),
    "sus2_tune_set": (_i, [_vp, _i, _i, _i]),
    "sus2_tune_messages": (_i, [_vp, _i, _i, _ip]),
    "sus2_tune_get": (_i, [_vp, _i, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "sus2_autotune": (_i, [_vp, _i, _i, _i] + [_vp] * 5 + [_i, _vp, ctypes.POINTER(_i), _dp]),
    "sus2_efs_grad": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "sus2_loss_grad": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _d, _d, _d, _dp, _vp,
                            _vp]),
    "sus2_train_lbfgs": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _d, _d, _d, _i, _i, _d,
                              _vp, _ip, _vp]),
    "sus2_pair_create": (_i, [_vp, _i, ctypes.POINTER(_vp)]),
    "sus2_pair_free": (None, [_vp]),
    "sus2_pair_comm_width": (_i, [_vp]),
    "sus2_pair_begin": (_i, [_vp, _i, _i, _vp, ctypes.c_int64, _vp, _vp, _vp]),
    "sus2_pair_pack_forward": (_i, [_vp, _i, _vp, _vp]),
    "sus2_pair_unpack_forward": (_i, [_vp, _i, _vp, _vp]),
    "sus2_pair_outer": (_i, [_vp]),
    "sus2_pair_pack_reverse": (_i, [_vp, _i, _vp, _vp]),
    "sus2_pair_add_reverse": (_i, [_vp, _i, _vp, _vp]),
    "sus2_pair_finish": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "sus2_profile_enable": (_i, [_vp, _i]),
    "sus2_profile_reset": (_i, [_vp]),
    "sus2_profile_read": (_i, [_vp, _i, ctypes.c_char_p, _dp, ctypes.POINTER(ctypes.c_int64), _ip]),
    "sus2_last_edge_count": (_i, [_vp, ctypes.POINTER(ctypes.c_int64)]),
}


class Sus2Error(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"[{STATUS.get(code, code)}] {message}")
        self.code = code


def lib():
    global _LIB
    with _LOCK:
        if _LIB is None:
            if not _PATH.exists():
                raise RuntimeError(f"{_PATH} is missing: build it with `python -m superlinear_amd.build` "
                                   "(there is no CPU fallback)")
            L = ctypes.CDLL(str(_PATH))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _LIB = L
    return _LIB


def check(code):
    if code != SUS2_OK:
        raise Sus2Error(code, lib().sus2_last_error().decode())


def library_path():
    return _PATH
