"""ctypes binding of oracle/liboracle.so (test infrastructure only).

Loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as
the checker; the product library never uses it.
"""
import ctypes as C
from pathlib import Path

HERE = Path(__file__).resolve().parent
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    path = HERE / "liboracle.so"
    if not path.exists():
        raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
    lib = C.CDLL(str(path))
    P = C.c_void_p
    D = C.POINTER(C.c_double)
    sig = {
        "oracle_read_ctl": (None, [C.c_int, P, P]),
        "oracle_scan_ctl": (C.c_double, [C.c_int, P, C.c_char_p, C.c_int, C.c_char_p, C.c_char_p]),
        "oracle_read_atm": (None, [C.c_char_p, C.c_char_p, P, P]),
        "oracle_read_obs": (None, [C.c_char_p, C.c_char_p, P, P]),
        "oracle_write_obs": (None, [C.c_char_p, C.c_char_p, P, P]),
        "oracle_read_aero": (None, [C.c_char_p, C.c_char_p, P, P]),
        "oracle_get_opt_prop": (None, [P, P]),
        "oracle_bhmie": (None, [C.c_double, C.c_double, C.c_double, D, D, D]),
        "oracle_gauher": (None, [D, D]),
        "oracle_formod": (None, [P, P, P, P]),
        "oracle_formod_pencil": (None, [P, P, P, P, C.c_int, C.c_int]),
        "oracle_raytrace": (None, [P, P, P, P, P, C.c_int]),
        "oracle_planck": (C.c_double, [C.c_double, C.c_double]),
        "oracle_brightness": (C.c_double, [C.c_double, C.c_double]),
        "oracle_suncoord": (None, [C.c_double, C.c_double, C.c_double, D, D]),
        "oracle_hydrostatic": (None, [P, P]),
        "oracle_intpol_atm_geo": (None, [P, P, C.c_double, C.c_double, C.c_double, D, D, D, D]),
        "oracle_continua": (None, [P, C.c_double, C.c_double, D, D, C.c_double, D, D]),
        "oracle_ctmn2": (C.c_double, [C.c_double, C.c_double, C.c_double]),
        "oracle_ctmo2": (C.c_double, [C.c_double, C.c_double, C.c_double]),
        "oracle_srcfunc_planck": (None, [P, C.c_double, D]),
        "oracle_formod_batch": (None, [P, C.c_int, C.c_int] + [D] * 12 + [C.c_int] * 3 +
                                [D] * 5 + [C.POINTER(C.c_int), P, C.c_int]),
        "oracle_atm2x": (C.c_size_t, [P, P, P, D, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "oracle_x2atm": (None, [P, D, P, P]),
        "oracle_obs2y": (C.c_size_t, [P, P, D, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "oracle_perturbation": (C.c_double, [P, C.c_int, C.c_double]),
        "oracle_kernel": (None, [P, P, P, P, D, C.c_size_t, C.c_size_t, C.c_size_t]),
        "oracle_kernel_columns": (None, [P, P, P, P, D, C.c_size_t, C.c_size_t, C.c_size_t,
                                         C.POINTER(C.c_int), C.c_int]),
        "oracle_set_fan": (None, [C.c_int]),
        "oracle_alpha_grid": (None, [C.c_int, D]),
        "oracle_set_mie_grid": (None, [C.c_int]),
        "oracle_bhmie180": (None, [C.c_double, C.c_double, C.c_double, D, D, D, D]),
        "oracle_read_ret": (None, [C.c_int, P, P, P]),
        "oracle_srcfunc_sca": (None, [P, P, P, C.c_double, D, D, C.c_int, D, C.c_int]),
        "oracle_cov_apr": (None, [P, P, P, P, D]),
        "oracle_cov_meas": (None, [P, P, P, D, D, D]),
        "oracle_optimal_estimation": (None, [P, P, P, P, P, P, P, P, D, C.POINTER(C.c_int), D]),
        "oracle_srcfunc_sca_sun": (None, [P, P, P, C.c_double, D, D, C.c_int, D]),
        "oracle_formod_fov": (None, [P, P]),
        "oracle_reset": (None, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib
