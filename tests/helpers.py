"""Shared test helpers: example cases and text comparison of obs files."""
import ctypes as C
from pathlib import Path

from pyjurassic import jtypes as J

# (control file, obs input, atmosphere, aerosol file or None, golden output)
EXAMPLE_CASES = {
    "clear": ("clear-air.ctl", "org/obs.tab", "atm_pwin_0_80.tab", None, "rad_clear.tab"),
    "ext1": ("extinction1.ctl", "org/obs.tab", "atm_midl_0_80.tab", "aero0.tab", "rad_ext1.tab"),
    "aero0": ("aerosol0.ctl", "org/obs_nadir.tab", "atm_pwin_0_60.tab", "aero0.tab", "rad_aero0.tab"),
    "aero1": ("aerosol1.ctl", "org/obs.tab", "atm_midl_0_80.tab", "aero1.tab", "rad_aero1.tab"),
}


def read_example(lib, prefix, case, keys=()):
    """Read ctl/obs/atm/aero of an example with `lib`'s readers (prefix
    'oracle_' for the oracle, '' for the GPU library); keys are extra
    KEY VALUE control arguments."""
    ctlf, obsf, atmf, aerof, _ = EXAMPLE_CASES[case]
    ctl, obs, atm, aero = J.ctl_t(), J.obs_t(), J.atm_t(), J.aero_t()
    extra = (["AEROFILE", aerof] if aerof else []) + list(keys)
    argc, argv = J.argv(ctlf, *extra)
    getattr(lib, prefix + "read_ctl")(argc, argv, C.byref(ctl))
    getattr(lib, prefix + "read_obs")(None, obsf.encode(), C.byref(ctl), C.byref(obs))
    getattr(lib, prefix + "read_atm")(None, atmf.encode(), C.byref(ctl), C.byref(atm))
    if aerof:
        getattr(lib, prefix + "read_aero")(None, aerof.encode(), C.byref(ctl), C.byref(aero))
        getattr(lib, prefix + "get_opt_prop")(C.byref(ctl), C.byref(aero))
    return ctl, obs, atm, aero


def obs_rows(path):
    """Numeric rows of an obs file (comment and blank lines skipped)."""
    rows = []
    for line in Path(path).read_text().splitlines():
        if line.strip() and not line.startswith("#"):
            rows.append([float(v) for v in line.split()])
    return rows


def read_workload(lib, prefix, w, ctlpath):
    """ctl (and aero with its optics, for aerosol workloads) of a synthetic
    workload, read with `lib`'s readers."""
    w.write_ctl(ctlpath)
    ctl = J.ctl_t()
    argc, argv = J.argv(ctlpath)
    getattr(lib, prefix + "read_ctl")(argc, argv, C.byref(ctl))
    aero = None
    if w.aerofile:
        aero = J.aero_t()
        getattr(lib, prefix + "read_aero")(None, w.aerofile.encode(), C.byref(ctl), C.byref(aero))
        getattr(lib, prefix + "get_opt_prop")(C.byref(ctl), C.byref(aero))
    return ctl, aero


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def oracle_batch(olib, w, ctlpath, nthreads=0):
    """The oracle's forward model over every ray of workload w: rad, tau
    [nray][nd], tangent point [3][nray], LOS points [nray]."""
    import numpy as np
    ctl, aero = read_workload(olib, "oracle_", w, ctlpath)
    n, nd = w.nray, ctl.nd
    rad, tau = np.zeros((n, nd)), np.zeros((n, nd))
    tp = np.zeros((3, n))
    npts = np.zeros(n, dtype=np.int32)
    g = [np.ascontiguousarray(w.geom[i], dtype=np.float64) for i in range(7)]
    # the oracle takes one atmosphere per ray
    rep = (lambda a: np.repeat(a, n, axis=0)) if w.natm == 1 else (lambda a: a)
    f = lambda a: np.ascontiguousarray(rep(a), dtype=np.float64)
    z, p, t, q = f(w.z), f(w.p), f(w.t), f(w.q)
    olib.oracle_formod_batch(C.byref(ctl), n, w.nlev, _dp(z), _dp(p), _dp(t), _dp(q), None,
                             *[_dp(x) for x in g], 0, n, nthreads, _dp(rad), _dp(tau),
                             _dp(tp[0]), _dp(tp[1]), _dp(tp[2]),
                             npts.ctypes.data_as(C.POINTER(C.c_int)),
                             C.byref(aero) if aero else None, ctl.sca_mult if aero else 0)
    return rad, tau, tp, npts


def gpu_batch(glib, w, ctlpath):
    """The GPU library's batched forward model (host pointers) over workload
    w; same outputs as oracle_batch (rad, tau as FP32)."""
    import numpy as np
    from pyjurassic import lib as L
    ctl, aero = read_workload(glib, "", w, ctlpath)
    n, nd = w.nray, ctl.nd
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    z, p, t, q = f(w.z), f(w.p), f(w.t), f(w.q)
    g = [f(w.geom[i]) for i in range(7)]
    rad = np.zeros((nd, n), dtype=np.float32)
    tau = np.zeros((nd, n), dtype=np.float32)
    tp = np.zeros((3, n))
    npts = np.zeros(n, dtype=np.int32)
    b = L.jur_batch_t()
    b.nray, b.nlev, b.natm = n, w.nlev, w.natm
    b.atm_of_ray = None
    b.z, b.p, b.t, b.q, b.k = z.ctypes.data, p.ctypes.data, t.ctypes.data, q.ctypes.data, None
    (b.time, b.obsz, b.obslon, b.obslat, b.vpz, b.vplon, b.vplat) = [x.ctypes.data for x in g]
    b.rad, b.tau = rad.ctypes.data, tau.ctypes.data
    b.tpz, b.tplon, b.tplat = tp[0].ctypes.data, tp[1].ctypes.data, tp[2].ctypes.data
    b.npts = npts.ctypes.data
    glib.jur_formod_batch(C.byref(ctl), C.byref(b), C.byref(aero) if aero else None,
                          ctl.sca_mult if aero else 0)
    return rad.T.astype(np.float64), tau.T.astype(np.float64), tp, npts
