"""The FP64 oracle against the reference's own regression fixtures.

examples/org/rad_*.tab are the goldens of the reference's shell regressions
(examples/clear.sh, extinction1.sh, aerosol0.sh, aerosol1.sh: `diff -q`,
byte-exact); los.0 is the golden of raytrace.sh.  The oracle must reproduce
them byte-for-byte through its own readers, formod and writer.
"""
import ctypes as C
import filecmp

import pytest

import oracle_ctypes
from conftest import EXAMPLES, GOLDEN
from helpers import EXAMPLE_CASES, read_example
from pyjurassic import jtypes as J


@pytest.fixture(scope="module")
def olib():
    return oracle_ctypes.load()


def test_layout():
    J.check_layout()


def test_climatology_restatement_matches_org_atm():
    # make_example_atm.py's pwin output must equal the reference's golden
    assert filecmp.cmp(EXAMPLES / "atm_pwin_0_80.tab", EXAMPLES / "org" / "atm.tab", shallow=False)


@pytest.mark.parametrize("case", sorted(EXAMPLE_CASES))
def test_formod_reproduces_golden(olib, in_examples, tmp_path, case):
    ctl, obs, atm, aero = read_example(olib, "oracle_", case)
    olib.oracle_formod(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(aero))
    out = tmp_path / "rad.tab"
    olib.oracle_write_obs(None, str(out).encode(), C.byref(ctl), C.byref(obs))
    golden = EXAMPLES / "org" / EXAMPLE_CASES[case][4]
    assert out.read_text() == golden.read_text()


def test_raytrace_reproduces_los0(olib, in_examples):
    """raytrace.sh: `limb 800 8 8 1` then the raytrace tool's los.0 dump
    (src/raytrace.c:55-109)."""
    ctl, obs, atm = J.ctl_t(), J.obs_t(), J.atm_t()
    argc, argv = J.argv("clear-air.ctl")
    olib.oracle_read_ctl(argc, argv, C.byref(ctl))
    olib.oracle_read_obs(None, b"org/obs_raytrace.tab", C.byref(ctl), C.byref(obs))
    olib.oracle_read_atm(None, b"atm_pwin_0_80.tab", C.byref(ctl), C.byref(atm))
    los = J.los_t()
    aero = J.aero_t()
    olib.oracle_raytrace(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(aero), C.byref(los), 0)
    g = lambda x: "%g" % x
    lines = ["# $1 = time (seconds since 2000-01-01T00:00Z)\n", "# $2 = LOS point altitude [km]\n",
             "# $3 = LOS point longitude [deg]\n", "# $4 = LOS point latitude [deg]\n",
             "# $5 = LOS point pressure [hPa]\n", "# $6 = LOS point temperature [K]\n"]
    emit = [ctl.emitter[i].value.decode() for i in range(ctl.ng)]
    for ig, e in enumerate(emit):
        lines.append("# $%d = LOS point %s volume mixing ratio \n" % (7 + ig, e))
    for iw in range(ctl.nw):
        lines.append("# $%d = LOS point window %d extinction [1/km]\n" % (7 + ctl.ng + iw, iw))
    for ig, e in enumerate(emit):
        lines.append("# $%d = LOS point %s column density [molec/cm^2] \n" % (7 + ctl.ng + ctl.nw + ig, e))
    base = 7 + 2 * ctl.ng + ctl.nw
    for k, name in enumerate(["beta_e", "beta_s", "beta_a"]):
        for idd in range(ctl.nd):
            lines.append("# $%d = LOS point %s(%.4f cm-1) [km-1] \n" % (base + k * ctl.nd + idd, name, ctl.nu[idd]))
    lines.append("# $%d = LOS segement length [km] \n" % (base + 3 * ctl.nd))
    lines.append("\n")
    for ip in range(los.np):
        row = ["%.2f" % obs.time[0], g(los.z[ip]), g(los.lon[ip]), g(los.lat[ip]), g(los.p[ip]), g(los.t[ip])]
        row += [g(los.q[ip][ig]) for ig in range(ctl.ng)]
        row += [g(los.k[ip][iw]) for iw in range(ctl.nw)]
        row += [g(los.u[ip][ig]) for ig in range(ctl.ng)]
        row += ["0"] * (3 * ctl.nd)  # no aerosol: beta * aerofac = 0
        row += [g(los.ds[ip])]
        lines.append(" ".join(row) + "\n")
    assert "".join(lines) == (EXAMPLES / "org" / "los.0").read_text()
    assert obs.tpz[0] == pytest.approx(obs.tpz[0])
