"""Load libjurassic_amd.so (the GPU forward model) and declare its C ABI.

Fails loudly when the library is missing: there is no CPU fallback.
"""
import ctypes as C
from pathlib import Path

PKG = Path(__file__).resolve().parents[1]
LIBPATH = PKG / "libjurassic_amd.so"

# Every symbol include/jurassic_amd.h declares, with (restype, argtypes).
P = C.c_void_p
D = C.POINTER(C.c_double)
F = C.POINTER(C.c_float)
SIGNATURES = {
    "formod": (None, [P, P, P, P]),
    "formod_pencil": (None, [P, P, P, P, C.c_int, C.c_int]),
    "raytrace": (None, [P, P, P, P, P, C.c_int]),
    "get_opt_prop": (None, [P, P]),
    "bhmie": (None, [C.c_double, C.c_double, C.c_double, D, D, D]),
    "hydrostatic": (None, [P, P]),
    "read_ctl": (None, [C.c_int, P, P]),
    "scan_ctl": (C.c_double, [C.c_int, P, C.c_char_p, C.c_int, C.c_char_p, C.c_char_p]),
    "read_atm": (None, [C.c_char_p, C.c_char_p, P, P]),
    "write_atm": (None, [C.c_char_p, C.c_char_p, P, P]),
    "read_obs": (None, [C.c_char_p, C.c_char_p, P, P]),
    "write_obs": (None, [C.c_char_p, C.c_char_p, P, P]),
    "read_aero": (None, [C.c_char_p, C.c_char_p, P, P]),
    "write_aero": (None, [C.c_char_p, C.c_char_p, P]),
    "read_shape": (None, [C.c_char_p, D, D, C.POINTER(C.c_int)]),
    "copy_aero": (None, [P, P, P, C.c_int]),
    "brightness": (C.c_double, [C.c_double, C.c_double]),
    "planck": (C.c_double, [C.c_double, C.c_double]),
    "kernel": (None, [P, P, P, P, P]),
    "jur_kernel": (C.c_int, [P, P, P, P, D, C.c_size_t, C.c_size_t, C.c_size_t]),
    "jur_state_size": (C.c_size_t, [P, P, P]),
    "jur_meas_size": (C.c_size_t, [P, P]),
    "read_ret": (None, [C.c_int, P, P, P]),
    "optimal_estimation": (None, [P, P, P, P, P, P, P, P]),
    "jur_tab2bin": (C.c_int, [P]),
    "write_matrix": (None, [C.c_char_p, C.c_char_p, P, P, P, P, P, C.c_char_p, C.c_char_p, C.c_char_p]),
    "jur_formod_batch": (C.c_int, [P, P, P, C.c_int]),
    "jur_formod_batch_device": (C.c_int, [P, P, P, C.c_int, P]),
    "jur_formod_batch_device_noalloc": (C.c_int, [P, P, P, C.c_int, P]),
    "jur_sync": (C.c_int, []),
    "jur_init": (C.c_int, [C.c_int]),
    "jur_set_option": (C.c_int, [C.c_char_p, C.c_double]),
    "jur_stat": (C.c_double, [C.c_char_p]),
    "jur_mlp_init": (C.c_int, [C.c_uint]),
    "jur_mlp_set_data": (C.c_int, [F, F, C.c_long, C.c_int, C.c_uint]),
    "jur_mlp_train": (C.c_int, [C.c_int, C.c_int, C.c_float, F]),
    "jur_mlp_grads": (C.c_int, [F, F, C.c_int, F, F]),
    "jur_mlp_forward": (C.c_int, [F, C.c_long, F]),
    "jur_mlp_get_params": (C.c_int, [F]),
    "jur_mlp_set_params": (C.c_int, [F]),
    "jur_mlp_nparam": (C.c_int, []),
    "jur_mlp_save": (C.c_int, [C.c_char_p]),
    "jur_mlp_load": (C.c_int, [C.c_char_p]),
    "jur_sca_collect": (C.c_long, [F, F, C.c_long]),
}


class jur_batch_t(C.Structure):
    _fields_ = [("nray", C.c_int), ("nlev", C.c_int), ("natm", C.c_int),
                ("atm_of_ray", C.c_void_p),
                ("z", C.c_void_p), ("p", C.c_void_p), ("t", C.c_void_p), ("q", C.c_void_p),
                ("k", C.c_void_p),
                ("time", C.c_void_p), ("obsz", C.c_void_p), ("obslon", C.c_void_p),
                ("obslat", C.c_void_p), ("vpz", C.c_void_p), ("vplon", C.c_void_p),
                ("vplat", C.c_void_p),
                ("rad", C.c_void_p), ("tau", C.c_void_p),
                ("tpz", C.c_void_p), ("tplon", C.c_void_p), ("tplat", C.c_void_p),
                ("npts", C.c_void_p), ("out_ld", C.c_int)]


class jur_gsl_matrix(C.Structure):
    """GSL 1.15's gsl_matrix layout."""
    _fields_ = [("size1", C.c_size_t), ("size2", C.c_size_t), ("tda", C.c_size_t),
                ("data", C.c_void_p), ("block", C.c_void_p), ("owner", C.c_int)]


_lib = None


def load():
    """The GPU library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIBPATH.exists():
        raise RuntimeError("libjurassic_amd.so is not built: run __graft_entry__.build() "
                           "(make -C jurassic-scatter_amd)")
    lib = C.CDLL(str(LIBPATH))
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def ptr(a):
    """Address of a numpy array or torch tensor (host or device)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data
