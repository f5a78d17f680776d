"""optimal_estimation (reference src/retrievalmodel.c:407-706) with the GPU
forward model and Jacobian against the oracle's restatement.

Same a priori, same measurements (the oracle's forward model of a perturbed
truth) on both sides.  The product's forward and Jacobian calls integrate in
FP64 (option kernel_fp64, default), so the iterations take the same
accept / reject path as the oracle's (each case checks that its decisions
are clear of rounding: near convergence with a large damping the cost
comparisons become ties that either path may break) and the retrieved
states agree to ~1e-10 K; the text outputs (cost file, error profiles: %g,
6 digits) agree to their print precision.
"""
import ctypes as C
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle_ctypes
from helpers import read_example
from pyjurassic import jtypes as J
from pyjurassic import lib as L

pytestmark = pytest.mark.gpu

BASE = ("RETT_ZMIN", 5, "RETT_ZMAX", 30, "ERR_TEMP", 5, "ERR_TEMP_CZ", 3, "ERR_TEMP_CH", 500,
        "ERR_NOISE[0]", 1e-6, "ERR_FORMOD[0]", 0.5, "CONV_DMIN", 1e-6)
CASES = {
    # (extra keys, temperature bump of the truth [K])
    "temperature": ((), 3.0),
    # large bump: rejected steps raise the damping; Jacobian every 2nd
    # iteration; six iterations (later ones compare costs equal to rounding)
    "rejections_recomp2": (("KERNEL_RECOMP", 2, "CONV_ITMAX", 6), 60.0),
    # temperature and H2O (correlated a priori), matrices written
    "t_h2o_matrices": (("RETQ_ZMIN[1]", 5, "RETQ_ZMAX[1]", 20, "ERR_Q[1]", 30, "ERR_Q_CZ[1]", 2,
                        "ERR_Q_CH[1]", 300, "WRITE_MATRIX", 1), 3.0),
    # extinction example: temperature and the aerosol layer's particle
    # concentration (Mie optics refreshed for every state), truth 1.5 N
    "t_aerosol_nn": (("RETNN", 1, "ERR_NN", 1.0, "CONV_ITMAX", 3), 3.0),
}
CASE_OF = {"t_aerosol_nn": "ext1"}


@pytest.fixture(scope="module")
def glib():
    return L.load()


@pytest.fixture(scope="module")
def olib():
    return oracle_ctypes.load()


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _copy(s):
    d = type(s)()
    C.memmove(C.addressof(d), C.addressof(s), C.sizeof(s))
    return d


def setup(lib, prefix, keys, case="clear"):
    ctl, obs, atm, aero = read_example(lib, prefix, case, keys)
    ret = J.ret_t()
    argc, argv = J.argv("clear-air.ctl", *keys)
    getattr(lib, prefix + "read_ret")(argc, argv, C.byref(ctl), C.byref(ret))
    return ctl, ret, obs, atm, aero


def measurements(olib, keys, bump, case="clear"):
    ctl, _, obs, atm, aero = setup(olib, "oracle_", keys, case)
    tru = _copy(atm)
    for ip in range(atm.np):
        tru.t[ip] += bump * np.exp(-((atm.z[ip] - 18) / 5) ** 2)
    if case != "clear":
        aero.nn[0] *= 1.5
        olib.oracle_get_opt_prop(C.byref(ctl), C.byref(aero))
    olib.oracle_formod(C.byref(ctl), C.byref(tru), C.byref(obs), C.byref(aero))
    return obs


def rows(path):
    return np.array([[float(v) for v in ln.split()] for ln in Path(path).read_text().splitlines()
                     if ln.strip() and not ln.startswith("#")])


@pytest.mark.parametrize("name", sorted(CASES))
def test_retrieval_vs_oracle(glib, olib, in_examples, tmp_path, name):
    extra, bump = CASES[name]
    case = CASE_OF.get(name, "clear")
    keys = BASE + extra
    print("case %s" % name, flush=True)
    meas = measurements(olib, keys, bump, case)
    # oracle
    ctl, ret, _, atm, aero = setup(olib, "oracle_", keys, case)
    n = olib.oracle_atm2x(C.byref(ctl), C.byref(atm), C.byref(aero), None, None, None)
    chisq, nc, err = np.zeros(1 + 20 * ret.conv_itmax), C.c_int(0), np.zeros(n)
    o_obs, o_atm, o_aero = J.obs_t(), J.atm_t(), J.aero_t()
    olib.oracle_optimal_estimation(C.byref(ret), C.byref(ctl), C.byref(_copy(meas)), C.byref(o_obs),
                                   C.byref(atm), C.byref(o_atm), C.byref(aero), C.byref(o_aero), _dp(chisq),
                                   C.byref(nc), _dp(err))
    chisq = chisq[:nc.value]
    # every accept / reject decision of the case is clear of rounding: the
    # two FP64 paths differ by ~1e-13, so both take the same steps
    old, gaps = chisq[0], []
    for v in chisq[1:]:
        gaps.append(abs(v - old) / old)
        old = min(old, v)
    assert min(gaps) > 1e-9, min(gaps)
    # GPU library
    ctl, ret, _, atm, aero = setup(glib, "", keys, case)
    ret.dir = str(tmp_path).encode()
    g_obs, g_atm, g_aero = J.obs_t(), J.atm_t(), J.aero_t()
    glib.optimal_estimation(C.byref(ret), C.byref(ctl), C.byref(_copy(meas)), C.byref(g_obs), C.byref(atm),
                            C.byref(g_atm), C.byref(aero), C.byref(g_aero))
    costs = rows(tmp_path / "costs.tab")
    print("%s: %d cost evaluations, final chi^2/m %.4g" % (name, len(chisq), chisq[-1]), flush=True)
    assert len(costs) == len(chisq)
    np.testing.assert_allclose(costs[:, 1], chisq, rtol=1e-5)
    assert np.all(costs[:, 2] + costs[:, 3] == pytest.approx(costs[:, 1], rel=2e-5))
    # retrieved state (in memory) and final radiances
    tg, to = np.array(g_atm.t[:g_atm.np]), np.array(o_atm.t[:o_atm.np])
    dt = np.max(np.abs(tg - to))
    print("%s: max |T_gpu - T_oracle| %.2e K" % (name, dt), flush=True)
    assert dt < 1e-5
    qg, qo = np.array(g_atm.q[1][:g_atm.np]), np.array(o_atm.q[1][:o_atm.np])
    np.testing.assert_allclose(qg, qo, rtol=1e-7)
    if "RETNN" in extra:
        print("%s: N a priori %g, retrieved gpu %.9g oracle %.9g" % (name, aero.nn[0], g_aero.nn[0], o_aero.nn[0]),
              flush=True)
        assert g_aero.nn[0] == pytest.approx(o_aero.nn[0], rel=1e-7)
        assert o_aero.nn[0] / aero.nn[0] - 1 > 0.005  # the concentration moved towards the truth
    rg = np.array(g_obs.rad[0][:g_obs.nr])
    ro = np.array(o_obs.rad[0][:o_obs.nr])
    np.testing.assert_allclose(rg, ro, rtol=1e-8)
    # total retrieval error of the temperature levels (atm_err_total.tab, column 6)
    z = np.array(atm.z[:atm.np])
    lt = np.where((z >= 5) & (z <= 30))[0]
    et = rows(tmp_path / "atm_err_total.tab")[lt, 5]
    np.testing.assert_allclose(et, err[:len(lt)], rtol=2e-5)
    for f in ("atm_final.tab", "obs_final.tab", "atm_err_noise.tab", "atm_err_formod.tab", "atm_cont.tab",
              "atm_res.tab"):
        assert (tmp_path / f).exists(), f
    if "WRITE_MATRIX" in extra:
        for f in ("matrix_cov_apr.tab", "matrix_kernel.tab", "matrix_cov_ret.tab", "matrix_gain.tab",
                  "matrix_avk.tab"):
            assert (tmp_path / f).stat().st_size > 0, f
        # degrees of freedom, the trace of the averaging kernel, lie between
        # 0 and the number of measurements; the file lists (i, j) pairs with
        # the element last and leaves out zeros
        avk = rows_matrix(tmp_path / "matrix_avk.tab")
        dof = sum(avk.get((i, i), 0.0) for i in range(n))
        print("%s: degrees of freedom %.3f of m = %d" % (name, dof, len(o_obs.rad[0][:o_obs.nr])), flush=True)
        assert 0 < dof <= o_obs.nr


def rows_matrix(path):
    out = {}
    for ln in Path(path).read_text().splitlines():
        if ln.strip() and not ln.startswith("#"):
            v = ln.split()
            out[(int(v[0]), int(v[6]))] = float(v[-1])
    return out


def test_kernel_recomp_must_be_positive(in_examples, tmp_path):
    code = (
        "import ctypes as C, sys; sys.path[:0] = %r\n"
        "import torch\n"
        "from pyjurassic import lib as L, jtypes as J; from helpers import read_example\n"
        "lib = L.load()\n"
        "keys = ('RETT_ZMIN', 5, 'RETT_ZMAX', 10, 'KERNEL_RECOMP', 0)\n"
        "ctl, obs, atm, aero = read_example(lib, '', 'clear', keys)\n"
        "ret = J.ret_t(); argc, argv = J.argv('clear-air.ctl', *keys)\n"
        "lib.read_ret(argc, argv, C.byref(ctl), C.byref(ret)); ret.dir = %r\n"
        "o, a, e = J.obs_t(), J.atm_t(), J.aero_t()\n"
        "lib.optimal_estimation(C.byref(ret), C.byref(ctl), C.byref(obs), C.byref(o), C.byref(atm),"
        " C.byref(a), C.byref(aero), C.byref(e))\n" % (sys.path[:4], str(tmp_path).encode()))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "KERNEL_RECOMP must be positive" in r.stdout
