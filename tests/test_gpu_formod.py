"""GPU forward model against the oracle and the reference fixtures.

Radiances are FP32 on the GPU (EGA, path integration with Kahan-compensated
radiance) and FP64 in the oracle; the bar is 1e-4 relative radiance
(<= 0.01 K brightness temperature) on every ray and channel, transmittance
within 1e-5 absolute, tangent points within 1e-6 km (FP64 geometry).
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_ctypes
from conftest import EXAMPLES
from helpers import EXAMPLE_CASES, obs_rows, read_example
from pyjurassic import jtypes as J
from pyjurassic import lib as L

pytestmark = pytest.mark.gpu

RAD_RTOL = 1e-4
TAU_ATOL = 1e-5


@pytest.fixture(scope="module")
def glib():
    return L.load()


@pytest.fixture(scope="module")
def olib():
    return oracle_ctypes.load()


def arrays(ctl, obs):
    nr, nd = obs.nr, ctl.nd
    rad = np.array([[obs.rad[i][r] for r in range(nr)] for i in range(nd)])
    tau = np.array([[obs.tau[i][r] for r in range(nr)] for i in range(nd)])
    tp = np.array([[obs.tpz[r], obs.tplon[r], obs.tplat[r]] for r in range(nr)])
    return rad, tau, tp


@pytest.mark.parametrize("case", ["clear", "ext1"])
def test_examples_vs_oracle(glib, olib, in_examples, tmp_path, case):
    ctl, obs, atm, aero = read_example(olib, "oracle_", case)
    olib.oracle_formod(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(aero))
    ref = arrays(ctl, obs)
    ctl2, obs2, atm2, aero2 = read_example(glib, "", case)
    glib.formod(C.byref(ctl2), C.byref(atm2), C.byref(obs2), C.byref(aero2))
    got = arrays(ctl2, obs2)
    print(case, "max rel rad", np.max(np.abs(got[0] / ref[0] - 1)), "max abs tau",
          np.max(np.abs(got[1] - ref[1])), flush=True)
    np.testing.assert_allclose(got[0], ref[0], rtol=RAD_RTOL)
    np.testing.assert_allclose(got[1], ref[1], atol=TAU_ATOL)
    np.testing.assert_allclose(got[2], ref[2], atol=1e-6)
    # and against the reference's own golden file (%g, 6 digits)
    out = tmp_path / "rad.tab"
    glib.write_obs(None, str(out).encode(), C.byref(ctl2), C.byref(obs2))
    g = np.array(obs_rows(EXAMPLES / "org" / EXAMPLE_CASES[case][4]))
    m = np.array(obs_rows(out))
    np.testing.assert_allclose(m[:, 10], g[:, 10], rtol=RAD_RTOL)
    np.testing.assert_allclose(m[:, 11], g[:, 11], atol=TAU_ATOL, rtol=1e-4)
    np.testing.assert_allclose(m[:, 7:10], g[:, 7:10], rtol=1e-5, atol=1e-6)


def test_raytrace_los_vs_oracle(glib, olib, in_examples):
    ctl, obs, atm = J.ctl_t(), J.obs_t(), J.atm_t()
    argc, argv = J.argv("clear-air.ctl")
    olib.oracle_read_ctl(argc, argv, C.byref(ctl))
    olib.oracle_read_obs(None, b"org/obs_raytrace.tab", C.byref(ctl), C.byref(obs))
    olib.oracle_read_atm(None, b"atm_pwin_0_80.tab", C.byref(ctl), C.byref(atm))
    aero = J.aero_t()
    lo, lg = J.los_t(), J.los_t()
    olib.oracle_raytrace(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(aero), C.byref(lo), 0)
    tpo = (obs.tpz[0], obs.tplon[0], obs.tplat[0])
    glib.raytrace(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(aero), C.byref(lg), 0)
    assert lg.np == lo.np == 216
    n = lo.np
    for f in ["z", "lon", "lat", "p", "t", "ds"]:
        a = np.frombuffer(getattr(lo, f), dtype=np.float64)[:n]
        b = np.frombuffer(getattr(lg, f), dtype=np.float64)[:n]
        np.testing.assert_allclose(b, a, rtol=1e-11, atol=1e-10, err_msg=f)
    uo = np.array([[lo.u[i][g] for g in range(ctl.ng)] for i in range(n)])
    ug = np.array([[lg.u[i][g] for g in range(ctl.ng)] for i in range(n)])
    np.testing.assert_allclose(ug, uo, rtol=1e-11)
    np.testing.assert_allclose((obs.tpz[0], obs.tplon[0], obs.tplat[0]), tpo, rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("case", ["aero0", "aero1"])
def test_scattering_examples_vs_oracle(glib, olib, in_examples, tmp_path, case):
    """Single scattering (SCA_MULT=1): nadir ice cloud (aerosol0.sh) and limb
    ice + H2SO4 layers (aerosol1.sh), against the oracle and the golden."""
    ctl, obs, atm, aero = read_example(olib, "oracle_", case)
    olib.oracle_formod(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(aero))
    ref = arrays(ctl, obs)
    ctl2, obs2, atm2, aero2 = read_example(glib, "", case)
    glib.formod(C.byref(ctl2), C.byref(atm2), C.byref(obs2), C.byref(aero2))
    got = arrays(ctl2, obs2)
    print(case, "max rel rad", np.max(np.abs(got[0] / ref[0] - 1)), "max abs tau",
          np.max(np.abs(got[1] - ref[1])), "pencils", glib.jur_stat(b"pencils"), flush=True)
    np.testing.assert_allclose(got[0], ref[0], rtol=RAD_RTOL)
    np.testing.assert_allclose(got[1], ref[1], atol=TAU_ATOL)
    out = tmp_path / "rad.tab"
    glib.write_obs(None, str(out).encode(), C.byref(ctl2), C.byref(obs2))
    g = np.array(obs_rows(EXAMPLES / "org" / EXAMPLE_CASES[case][4]))
    m = np.array(obs_rows(out))
    np.testing.assert_allclose(m[:, 10], g[:, 10], rtol=RAD_RTOL)
    np.testing.assert_allclose(m[:, 11], g[:, 11], atol=TAU_ATOL, rtol=1e-4)


def test_mie_layer_optics_vs_oracle(glib, olib, in_examples):
    """get_opt_prop on the device (FP64 Mie) against the oracle: the six modes
    in three layers of aero1.tab."""
    ctl, _, _, aero = read_example(olib, "oracle_", "aero1")
    ctl2, _, _, aero2 = read_example(glib, "", "aero1")
    assert aero2.nl == aero.nl == 3
    for il in range(aero.nl):
        for f in ["beta_e", "beta_s", "beta_a"]:
            a = getattr(aero, f)[il][0]
            b = getattr(aero2, f)[il][0]
            assert b == pytest.approx(a, rel=1e-12), (f, il)
        pa = np.array(aero.p[il][0])
        pb = np.array(aero2.p[il][0])
        np.testing.assert_allclose(pb, pa, rtol=1e-11)


def test_mie_grid_layer_optics_vs_oracle(glib, olib, in_examples):
    """Option mie_grid (the aerosol configuration's 64 log-spaced radii and
    180 phase angles) against the oracle's mie_grid on the six modes of
    aero1.tab: the device's 180-angle series, resampling and radius rule."""
    olib.oracle_set_mie_grid(1)
    glib.jur_set_option(b"mie_grid", 1)
    try:
        ctl, _, _, aero = read_example(olib, "oracle_", "aero1")
        ctl2, _, _, aero2 = read_example(glib, "", "aero1")
    finally:
        olib.oracle_set_mie_grid(0)
        glib.jur_set_option(b"mie_grid", 0)
    assert glib.jur_stat(b"mie_spheres") > 0
    for il in range(aero.nl):
        for f in ["beta_e", "beta_s", "beta_a"]:
            assert getattr(aero2, f)[il][0] == pytest.approx(getattr(aero, f)[il][0], rel=1e-12), (f, il)
        np.testing.assert_allclose(np.array(aero2.p[il][0]), np.array(aero.p[il][0]), rtol=1e-11)


def test_bhmie_vs_oracle(glib, olib):
    for x, nr, ni in [(0.05, 1.33, 0.0), (2.5, 1.5, 0.01), (40.0, 1.31, 0.3), (785.0, 1.45, 0.2)]:
        pa, pb = (C.c_double * 181)(), (C.c_double * 181)()
        qa, sa, qb, sb = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        olib.oracle_bhmie(x, nr, ni, pa, C.byref(qa), C.byref(sa))
        glib.bhmie(x, nr, ni, pb, C.byref(qb), C.byref(sb))
        assert qb.value == pytest.approx(qa.value, rel=1e-11)
        assert sb.value == pytest.approx(sa.value, rel=1e-11)
        np.testing.assert_allclose(np.array(pb), np.array(pa), rtol=1e-10)


def _bin_table(path):
    """The reference's binary table layout (forwardmodel.c:636-651)."""
    b = path.read_bytes()
    o = 0

    def take(fmt, n):
        nonlocal o
        a = np.frombuffer(b, dtype=fmt, count=n, offset=o)
        o += a.nbytes
        return a
    np_ = int(take("<i4", 1)[0])
    p = take("<f8", np_)
    nt = take("<i4", np_)
    cols = []
    for ip in range(np_):
        t = take("<f8", nt[ip])
        nu = take("<i4", nt[ip])
        for it in range(nt[ip]):
            u = take("<f4", nu[it])
            e = take("<f4", nu[it])
            cols.append((p[ip], t[it], u, e))
    return cols


def test_ascii_tables_vs_oracle(glib, olib, in_examples, tmp_path):
    """The ASCII table reader (forwardmodel.c:658-737): the clear example's
    binary tables written as text, each (p, T) column with one injected
    non-increasing row (which overwrites the column's last node, as in the
    reference).  Both libraries read the text; formod agrees with the oracle,
    and without the injected rows the text tables give the binary tables'
    radiances exactly."""
    import shutil
    out = {}
    for inject in (False, True):
        d = tmp_path / ("inj" if inject else "plain")
        d.mkdir()
        for g in ("CO2", "H2O", "O3"):
            lines = []
            for p, t, u, e in _bin_table(EXAMPLES / "tab" / ("boxcar_792.0000_%s.bin" % g)):
                for k in range(len(u)):
                    lines.append("%.9g %.9g %.9g %.9g" % (p, t, u[k], e[k]))
                    if inject and k == len(u) // 2:
                        lines.append("%.9g %.9g %.9g %.9g" % (p, t, u[k] * 0.5, e[k] * 0.5))
            (d / ("boxcar_792.0000_%s.tab" % g)).write_text("\n".join(lines) + "\n")
        shutil.copy(EXAMPLES / "tab" / "boxcar_792.0000.filt", d / "boxcar_792.0000.filt")
        keys = ("TBLBASE", str(d / "boxcar"))
        for lib, prefix in ((olib, "oracle_"), (glib, "")):
            ctl, obs, atm = J.ctl_t(), J.obs_t(), J.atm_t()
            argc, argv = J.argv("clear-air.ctl", *keys)
            getattr(lib, prefix + "read_ctl")(argc, argv, C.byref(ctl))
            getattr(lib, prefix + "read_obs")(None, b"org/obs.tab", C.byref(ctl), C.byref(obs))
            getattr(lib, prefix + "read_atm")(None, b"atm_pwin_0_80.tab", C.byref(ctl), C.byref(atm))
            getattr(lib, prefix + "formod")(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(J.aero_t()))
            out[(inject, prefix)] = np.array([obs.rad[0][i] for i in range(obs.nr)])
    ctl, obs, atm = J.ctl_t(), J.obs_t(), J.atm_t()
    argc, argv = J.argv("clear-air.ctl")
    glib.read_ctl(argc, argv, C.byref(ctl))
    glib.read_obs(None, b"org/obs.tab", C.byref(ctl), C.byref(obs))
    glib.read_atm(None, b"atm_pwin_0_80.tab", C.byref(ctl), C.byref(atm))
    glib.formod(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(J.aero_t()))
    binary = np.array([obs.rad[0][i] for i in range(obs.nr)])
    np.testing.assert_array_equal(out[(False, "")], binary)
    for inject in (False, True):
        rel = np.max(np.abs(out[(inject, "")] / out[(inject, "oracle_")] - 1))
        print("ascii tables (injected rows %s): max rel rad vs oracle %.2e" % (inject, rel), flush=True)
        assert rel < 1e-4
    assert np.max(np.abs(out[(True, "oracle_")] / out[(False, "oracle_")] - 1)) > 1e-6  # the overwrite matters


@pytest.mark.parametrize("case,depth", [("clear", 0), ("ext1", 0), ("aero0", 0), ("aero0", 1), ("aero1", 1)])
def test_formod_pencil_vs_oracle(glib, olib, in_examples, case, depth):
    """The exported formod_pencil (reference src/forwardmodel.h:32-37,
    forwardmodel.c:224-405), called ray by ray at scattering depth 0
    (extinction only with aerosol) and 1, against the oracle's pencil:
    only ray ir's radiance, transmittance and tangent point change."""
    ctl, obs, atm, aero = read_example(olib, "oracle_", case)
    ctl2, obs2, atm2, aero2 = read_example(glib, "", case)
    rays = range(0, obs.nr, max(1, obs.nr // 4))
    for ir in rays:
        olib.oracle_formod_pencil(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(aero), depth, ir)
        before = arrays(ctl2, obs2)
        glib.formod_pencil(C.byref(ctl2), C.byref(atm2), C.byref(obs2), C.byref(aero2), depth, ir)
        after = arrays(ctl2, obs2)
        others = [r for r in range(obs.nr) if r != ir]
        for k in (0, 1):  # rad, tau [nd][nr]
            np.testing.assert_array_equal(after[k][:, others], before[k][:, others])
        np.testing.assert_array_equal(after[2][others], before[2][others])  # tangent points [nr][3]
    sel = list(rays)
    ref, got = arrays(ctl, obs), arrays(ctl2, obs2)
    rel = np.abs(got[0][:, sel] / ref[0][:, sel] - 1)
    print("%s depth %d: %d pencils, max rel rad %.2e" % (case, depth, len(sel), rel.max()), flush=True)
    np.testing.assert_allclose(got[0][:, sel], ref[0][:, sel], rtol=RAD_RTOL)
    np.testing.assert_allclose(got[1][:, sel], ref[1][:, sel], atol=TAU_ATOL)
    np.testing.assert_allclose(got[2][sel], ref[2][sel], atol=1e-6)
