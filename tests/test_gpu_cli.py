"""The formod command (jurassic-scatter_amd/bin/formod, the reference CLI's
arguments and keys) on the reference examples: output files against the
goldens, as the reference's own regression scripts compare them, and
TASK c (per-emitter contributions and extinction only) against the oracle."""
import ctypes as C
import subprocess

import numpy as np
import pytest

import oracle_ctypes
from conftest import EXAMPLES, ROOT
from helpers import EXAMPLE_CASES, obs_rows
from pyjurassic import jtypes as J

pytestmark = pytest.mark.gpu

FORMOD = ROOT / "jurassic-scatter_amd" / "bin" / "formod"


def run_cli(args, tmp_path):
    r = subprocess.run([str(FORMOD)] + [str(a) for a in args], cwd=EXAMPLES, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout[-500:] + r.stderr[-500:]
    return r


@pytest.mark.parametrize("case", sorted(EXAMPLE_CASES))
def test_cli_vs_golden(case, tmp_path):
    ctlf, obsf, atmf, aerof, golden = EXAMPLE_CASES[case]
    out = tmp_path / "rad.tab"
    args = [ctlf, obsf, atmf, out] + (["AEROFILE", aerof] if aerof else [])
    run_cli(args, tmp_path)
    g = np.array(obs_rows(EXAMPLES / "org" / golden))
    m = np.array(obs_rows(out))
    assert m.shape == g.shape
    np.testing.assert_allclose(m[:, :10], g[:, :10], rtol=1e-5, atol=1e-6)  # geometry, tangent points
    np.testing.assert_allclose(m[:, 10], g[:, 10], rtol=1e-4)             # radiance
    np.testing.assert_allclose(m[:, 11], g[:, 11], rtol=1e-4, atol=1e-5)  # transmittance


def test_cli_task_c(tmp_path):
    """TASK c writes rad.tab plus rad.tab.<emitter> (only that emitter, no
    extinction, no N2/O2 continua) and rad.tab.EXTINCT (no emitters)."""
    out = tmp_path / "rad.tab"
    run_cli(["clear-air.ctl", "org/obs.tab", "atm_pwin_0_80.tab", out, "TASK", "c"], tmp_path)
    olib = oracle_ctypes.load()
    import os
    cwd = os.getcwd()
    os.chdir(EXAMPLES)
    try:
        ctl, obs, atm = J.ctl_t(), J.obs_t(), J.atm_t()
        argc, argv = J.argv("clear-air.ctl")
        olib.oracle_read_ctl(argc, argv, C.byref(ctl))
        olib.oracle_read_obs(None, b"org/obs.tab", C.byref(ctl), C.byref(obs))
        olib.oracle_read_atm(None, b"atm_pwin_0_80.tab", C.byref(ctl), C.byref(atm))
        ctl.ctm_n2 = 0
        ctl.ctm_o2 = 0
        aero = J.aero_t()
        names = [ctl.emitter[ig].value.decode() for ig in range(ctl.ng)]
        for ig, name in enumerate(names + ["EXTINCT"]):
            a2 = J.atm_t.from_buffer_copy(atm)
            for iw in range(ctl.nw):
                for ip in range(a2.np):
                    a2.k[iw][ip] = 0.0 if name != "EXTINCT" else a2.k[iw][ip]
            for ig2 in range(ctl.ng):
                if ig2 != ig:
                    for ip in range(a2.np):
                        a2.q[ig2][ip] = 0.0
            o2 = J.obs_t.from_buffer_copy(obs)
            olib.oracle_formod(C.byref(ctl), C.byref(a2), C.byref(o2), C.byref(aero))
            ref = np.array([o2.rad[0][i] for i in range(o2.nr)])
            got = np.array(obs_rows(tmp_path / ("rad.tab." + name)))[:, 10]
            # EXTINCT of this example is zero up to rounding (no extinction, no
            # emitters): 1e-9 is 1e-7 of the example's radiances
            np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-9, err_msg=name)
    finally:
        os.chdir(cwd)


RETRIEVAL = ROOT / "jurassic-scatter_amd" / "bin" / "retrieval"
RET_KEYS = ("RETT_ZMIN", 5, "RETT_ZMAX", 30, "ERR_TEMP", 5, "ERR_TEMP_CZ", 3, "ERR_TEMP_CH", 500,
            "ERR_FORMOD[0]", 1, "CONV_DMIN", 1e-6, "CONV_ITMAX", 6, "RESMAX", 10)


def _oracle_retrieve(olib, ctl, ret, meas, atm, aero):
    o_obs, o_atm, o_aero = J.obs_t(), J.atm_t(), J.aero_t()
    n = olib.oracle_atm2x(C.byref(ctl), C.byref(atm), C.byref(aero), None, None, None)
    chisq, nc = np.zeros(1 + 20 * ret.conv_itmax), C.c_int(0)
    olib.oracle_optimal_estimation(C.byref(ret), C.byref(ctl), C.byref(meas), C.byref(o_obs), C.byref(atm),
                                   C.byref(o_atm), C.byref(aero), C.byref(o_aero),
                                   chisq.ctypes.data_as(C.POINTER(C.c_double)), C.byref(nc), None)
    return o_obs, o_atm, n


def test_cli_retrieval_with_screening(tmp_path, monkeypatch):
    """The retrieval command (reference src/retrieval.c) on the extinction
    example: measurements of a 3 K warmer stratosphere, one of them spoiled
    by 50 %; RESMAX 10 screens it out and the retrieval runs again.  The
    oracle runs the same two passes; the final atmospheres agree to the
    files' print precision."""
    monkeypatch.chdir(EXAMPLES)
    olib = oracle_ctypes.load()
    d = tmp_path / "case"
    d.mkdir()
    ctl, obs, atm, aero = J.ctl_t(), J.obs_t(), J.atm_t(), J.aero_t()
    argc, argv = J.argv("extinction1.ctl", *RET_KEYS)
    olib.oracle_read_ctl(argc, argv, C.byref(ctl))
    ret = J.ret_t()
    olib.oracle_read_ret(argc, argv, C.byref(ctl), C.byref(ret))
    olib.oracle_read_obs(None, b"org/obs.tab", C.byref(ctl), C.byref(obs))
    olib.oracle_read_atm(None, b"atm_midl_0_80.tab", C.byref(ctl), C.byref(atm))
    olib.oracle_read_aero(None, b"aero0.tab", C.byref(ctl), C.byref(aero))
    olib.oracle_get_opt_prop(C.byref(ctl), C.byref(aero))
    # inputs of the directory: a priori = the example, measurements of a truth
    (d / "atm_apr.tab").write_text((EXAMPLES / "atm_midl_0_80.tab").read_text())
    (d / "aero_apr.tab").write_text((EXAMPLES / "aero0.tab").read_text())
    tru = J.atm_t()
    C.memmove(C.addressof(tru), C.addressof(atm), C.sizeof(atm))
    for ip in range(tru.np):
        tru.t[ip] += 3.0 * np.exp(-((tru.z[ip] - 20) / 6) ** 2)
    meas = J.obs_t()
    C.memmove(C.addressof(meas), C.addressof(obs), C.sizeof(obs))
    olib.oracle_formod(C.byref(ctl), C.byref(tru), C.byref(meas), C.byref(aero))
    meas.rad[0][3] *= 1.5
    olib.oracle_write_obs(str(d).encode(), b"obs_meas.tab", C.byref(ctl), C.byref(meas))
    (tmp_path / "dirlist").write_text(str(d) + "\n")
    # the command
    r = subprocess.run([str(RETRIEVAL), "extinction1.ctl", str(tmp_path / "dirlist")] + [str(k) for k in RET_KEYS],
                       cwd=EXAMPLES, capture_output=True, text=True, timeout=150)
    assert r.returncode == 0, r.stdout[-800:] + r.stderr[-500:]
    assert "Found 1 bad measurements. Redo retrieval..." in r.stdout and "Retrieval done..." in r.stdout
    # the oracle: the same file round trip, two passes
    m2, a2, ae2 = J.obs_t(), J.atm_t(), J.aero_t()
    olib.oracle_read_obs(str(d).encode(), b"obs_meas.tab", C.byref(ctl), C.byref(m2))
    olib.oracle_read_atm(str(d).encode(), b"atm_apr.tab", C.byref(ctl), C.byref(a2))
    olib.oracle_read_aero(str(d).encode(), b"aero_apr.tab", C.byref(ctl), C.byref(ae2))
    o_obs, o_atm, n = _oracle_retrieve(olib, ctl, ret, m2, a2, ae2)
    bad = [ir for ir in range(m2.nr)
           if abs(1 - o_obs.rad[0][ir] / m2.rad[0][ir]) >= ret.resmax / 100]
    assert bad == [3]
    m2.rad[0][3] = float("nan")
    o_obs, o_atm, n = _oracle_retrieve(olib, ctl, ret, m2, a2, ae2)
    got = np.array(obs_rows(d / "atm_final.tab"))
    want_t = np.array(o_atm.t[:o_atm.np])
    # %g keeps 6 digits: 3 decimals for temperatures of 100-999 K
    np.testing.assert_allclose(got[:, 5], want_t, rtol=0, atol=5e-4 + 1e-9)
    final = np.array(obs_rows(d / "obs_final.tab"))
    assert np.isnan(final[3, 10]) and np.all(np.isfinite(np.delete(final[:, 10], 3)))
    for f in ("costs.tab", "atm_err_total.tab", "atm_cont.tab", "atm_res.tab"):
        assert (d / f).stat().st_size > 0, f


KERNEL = ROOT / "jurassic-scatter_amd" / "bin" / "kernel"


def test_cli_kernel_vs_oracle(tmp_path):
    """The kernel command (reference src/kernel.c): the Jacobian of the clear
    example for temperature and H2O written by write_matrix (non-zero
    elements, measurement space along rows), against the oracle's kernel;
    without WRITE_MATRIX no file is written, as in the reference."""
    keys = ["RETT_ZMIN", 0, "RETT_ZMAX", 40, "RETQ_ZMIN[1]", 0, "RETQ_ZMAX[1]", 20]
    out = tmp_path / "kernel.tab"
    r = subprocess.run([str(KERNEL), "clear-air.ctl", "org/obs.tab", "atm_pwin_0_80.tab", "-", str(out)] +
                       [str(k) for k in keys], cwd=EXAMPLES, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and not out.exists(), r.stdout[-500:]
    r = subprocess.run([str(KERNEL), "clear-air.ctl", "org/obs.tab", "atm_pwin_0_80.tab", "-", str(out),
                        "WRITE_MATRIX", "1"] + [str(k) for k in keys], cwd=EXAMPLES, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-500:] + r.stderr[-500:]
    got = {}
    names = set()
    for ln in out.read_text().splitlines():
        if ln.strip() and not ln.startswith("#"):
            v = ln.split()
            got[(int(v[0]), int(v[6]))] = float(v[-1])
            names.add(v[7])
    assert names == {"TEMPERATURE", "H2O"}
    olib = oracle_ctypes.load()
    import os
    cwd = os.getcwd()
    os.chdir(EXAMPLES)
    try:
        ctl, obs, atm, aero = J.ctl_t(), J.obs_t(), J.atm_t(), J.aero_t()
        argc, argv = J.argv("clear-air.ctl", *keys)
        olib.oracle_read_ctl(argc, argv, C.byref(ctl))
        olib.oracle_read_obs(None, b"org/obs.tab", C.byref(ctl), C.byref(obs))
        olib.oracle_read_atm(None, b"atm_pwin_0_80.tab", C.byref(ctl), C.byref(atm))
        n = olib.oracle_atm2x(C.byref(ctl), C.byref(atm), C.byref(aero), None, None, None)
        m = olib.oracle_obs2y(C.byref(ctl), C.byref(obs), None, None, None)
        k = np.zeros((m, n))
        olib.oracle_kernel(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(aero),
                           k.ctypes.data_as(C.POINTER(C.c_double)), m, n, n)
    finally:
        os.chdir(cwd)
    kg = np.zeros((m, n))
    for (i, j), v in got.items():
        kg[i, j] = v
    assert set(got) == {(i, j) for i in range(m) for j in range(n) if k[i, j] != 0}
    # %g prints 6 significant digits
    np.testing.assert_allclose(kg, k, rtol=1e-5, atol=1e-9 * np.abs(k).max())


RAYTRACE = ROOT / "jurassic-scatter_amd" / "bin" / "raytrace"


def test_cli_raytrace_golden(tmp_path):
    """raytrace.sh: the raytrace command's los.0 on the example's limb
    geometry against the reference's org/los.0 (the script's `diff -q`).
    The ray lies in the meridian plane of longitude 0, so the longitude
    column is atan2 of a rounding-level Cartesian component (|lon| < 2e-7
    deg): its printed digits follow the last FP64 bits of that component,
    which fused multiply-adds and the device's atan2 need not share with
    the C library's, and it is compared within 1e-9 deg.  Every other
    column must match the golden's text, except numbers within 1e-5
    relative / 1e-9 absolute on at most 3 lines."""
    import shutil
    for f in ("clear-air.ctl", "atm_pwin_0_80.tab"):
        shutil.copy(EXAMPLES / f, tmp_path / f)
    (tmp_path / "org").mkdir()
    shutil.copy(EXAMPLES / "org" / "obs_raytrace.tab", tmp_path / "org" / "obs_raytrace.tab")
    r = subprocess.run([str(RAYTRACE), "clear-air.ctl", "org/obs_raytrace.tab", "atm_pwin_0_80.tab",
                        "TBLBASE", str(EXAMPLES / "tab" / "boxcar")],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-500:] + r.stderr[-500:]
    got = (tmp_path / "los.0").read_text().splitlines()
    want = (EXAMPLES / "org" / "los.0").read_text().splitlines()
    assert len(got) == len(want)
    LON = 2  # time, z, lon, lat, ...
    differ, nlon = [], 0
    for g, w in zip(got, want):
        if g.startswith("#") or not g.strip():
            assert g == w
            continue
        gs, ws = g.split(), w.split()
        assert len(gs) == len(ws)
        assert abs(float(ws[LON])) < 1e-6, "the golden's rays are no longer in the meridian plane"
        assert abs(float(gs[LON]) - float(ws[LON])) <= 1e-9
        nlon += gs[LON] != ws[LON]
        if gs[:LON] + gs[LON + 1:] != ws[:LON] + ws[LON + 1:]:
            differ.append((gs, ws))
    print("los.0: %d of %d lines differ in text outside the longitude (%d longitudes differ within 1e-9 deg)" %
          (len(differ), len(want), nlon), flush=True)
    assert len(differ) <= 3
    for g, w in differ:
        np.testing.assert_allclose(np.array(g, dtype=float), np.array(w, dtype=float), rtol=1e-5, atol=1e-9)
