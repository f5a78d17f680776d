"""The oracle's 2-D and 3-D atmosphere interpolation (reference
src/atmosphere.c:269-483, IP = 2 and 3) and the ray tracing through them.

The reference examples use IP = 1 only, so these cases are pinned by closed
forms computed here independently of the oracle's code:
  * IP 2 weights two profiles by the projection of the point onto the chord
    between them, computed here with a dot product (the reference uses the
    law of cosines), clamped to the nearer profile beyond either end;
  * IP 2 with identical profiles reproduces IP 1;
  * IP 3 with a symmetric vertical window reproduces a linear profile, and
    two columns mix with weights 1 - d/CX, d the surface chord
    2 RE sin(angle / 2);
  * IP 3 with no point in range gives NaN (0 / 0), as in the reference.
"""
import ctypes as C
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np
import pytest

import oracle_ctypes
from pyjurassic import jtypes as J

RE = 6367.421


@pytest.fixture(scope="module")
def olib():
    return oracle_ctypes.load()


def ctl_ip(olib, ip, cz=0.0, cx=0.0):
    """A control with two emitters and one channel (no tables are read)."""
    d = Path(tempfile.mkdtemp(prefix="jur_geo_"))
    (d / "geo.ctl").write_text("NG = 2\nEMITTER[0] = CO2\nEMITTER[1] = H2O\nND = 1\nNU[0] = 792\n")
    ctl = J.ctl_t()
    argc, argv = J.argv(d / "geo.ctl", "IP", ip, "CZ", cz, "CX", cx)
    olib.oracle_read_ctl(argc, argv, C.byref(ctl))
    assert ctl.ip == ip and ctl.ng == 2
    return ctl


def column_atm(ctl, lons, lats, toff, nz=41):
    """Columns at (lon, lat) with levels 0..nz-1 km, T = 250 - 2 z + toff[i],
    p = 1000 exp(-z / 7), H2O linear in z."""
    atm = J.atm_t()
    n = 0
    for lon, lat, dt in zip(lons, lats, toff):
        for l in range(nz):
            atm.z[n], atm.lon[n], atm.lat[n] = l, lon, lat
            atm.t[n] = 250.0 - 2.0 * l + dt
            atm.p[n] = 1000.0 * np.exp(-l / 7.0)
            for g in range(ctl.ng):
                atm.q[g][n] = 1e-4 * (1 + 0.01 * l) * (g + 1)
            n += 1
    atm.np = n
    return atm


def at(olib, ctl, atm, z, lon, lat):
    p, t = C.c_double(), C.c_double()
    q, k = np.zeros(J.NGMAX), np.zeros(J.NWMAX)
    olib.oracle_intpol_atm_geo(C.byref(ctl), C.byref(atm), z, lon, lat, C.byref(p), C.byref(t),
                               q.ctypes.data_as(C.POINTER(C.c_double)), k.ctypes.data_as(C.POINTER(C.c_double)))
    return p.value, t.value, q


def cart(lon, lat):
    lo, la = np.radians(lon), np.radians(lat)
    return RE * np.array([np.cos(la) * np.cos(lo), np.cos(la) * np.sin(lo), np.sin(la)])


def test_ip2_projection_weights(olib):
    ctl = ctl_ip(olib, 2)
    lons = [0.0, 2.0, 4.0]
    atm = column_atm(ctl, lons, [0.0] * 3, [0.0, 6.0, 12.0])
    for z, lon, lat, (ia, ib) in [(10.0, 1.0, 0.0, (0, 1)), (12.3, 0.3, 0.5, (0, 1)), (7.5, 3.2, -0.4, (2, 1)),
                                  (20.0, 2.9, 0.2, (1, 2))]:
        xa, xb, x0 = cart(lons[ia], 0.0), cart(lons[ib], 0.0), cart(lon, lat)
        r = np.clip(np.dot(x0 - xa, xb - xa) / np.dot(xb - xa, xb - xa), 0, 1)
        t_want = 250.0 - 2.0 * z + (1 - r) * 6.0 * ia + r * 6.0 * ib
        p, t, q = at(olib, ctl, atm, z, lon, lat)
        assert t == pytest.approx(t_want, abs=1e-9), (z, lon, lat)
        # pressure: exponential interpolation within each profile, identical
        # profiles in p here
        assert p == pytest.approx(1000.0 * np.exp(-z / 7.0), rel=1e-12)
        assert q[1] == pytest.approx(2e-4 * (1 + 0.01 * z), rel=1e-12)


def test_ip2_clamps_beyond_the_ends(olib):
    ctl = ctl_ip(olib, 2)
    atm = column_atm(ctl, [0.0, 2.0, 4.0], [0.0] * 3, [0.0, 6.0, 12.0])
    assert at(olib, ctl, atm, 5.0, 5.5, 0.0)[1] == pytest.approx(250.0 - 10.0 + 12.0, abs=1e-12)
    assert at(olib, ctl, atm, 5.0, -1.0, 0.3)[1] == pytest.approx(250.0 - 10.0, abs=1e-12)


def test_ip2_identical_profiles_equal_ip1(olib):
    ctl2, ctl1 = ctl_ip(olib, 2), ctl_ip(olib, 1)
    atm2 = column_atm(ctl2, [0.0, 3.0], [0.0, 1.0], [0.0, 0.0])
    atm1 = column_atm(ctl1, [0.0], [0.0], [0.0])
    for z in (0.0, 3.7, 17.25, 39.9):
        a, b = at(olib, ctl2, atm2, z, 1.3, 0.2), at(olib, ctl1, atm1, z, 1.3, 0.2)
        assert a[0] == pytest.approx(b[0], rel=1e-14) and a[1] == pytest.approx(b[1], rel=1e-14)


def test_ip3_symmetric_window_and_two_columns(olib):
    cx = 300.0
    ctl = ctl_ip(olib, 3, cz=2.5, cx=cx)
    atm = column_atm(ctl, [0.0], [0.0], [0.0])
    # levels 1 km apart inside +-2.5 km: weights 0.2, 0.6, 1, 0.6, 0.2 about
    # a level, so a linear T is reproduced
    for z in (5.0, 20.0):
        assert at(olib, ctl, atm, z, 0.0, 0.0)[1] == pytest.approx(250.0 - 2.0 * z, abs=1e-12)
    # two columns 1 degree apart on the equator, the second 10 K warmer
    atm = column_atm(ctl, [0.0, 1.0], [0.0, 0.0], [0.0, 10.0])
    d = 2 * RE * np.sin(np.radians(0.5))
    wb = 1 - d / cx
    t = at(olib, ctl, atm, 12.0, 0.0, 0.0)[1]
    assert t == pytest.approx(250.0 - 24.0 + 10.0 * wb / (1 + wb), abs=1e-9)
    # midway: equal horizontal weights
    t = at(olib, ctl, atm, 12.0, 0.5, 0.0)[1]
    assert t == pytest.approx(250.0 - 24.0 + 5.0, abs=1e-9)


def test_ip3_empty_neighbourhood_is_nan(olib):
    ctl = ctl_ip(olib, 3, cz=0.4, cx=50.0)
    atm = column_atm(ctl, [0.0], [0.0], [0.0])
    assert np.isnan(at(olib, ctl, atm, 10.0, 5.0, 0.0)[1])


def test_ip2_single_point_profile_exits(olib, tmp_path):
    code = (
        "import ctypes as C, sys; sys.path[:0] = %r\n"
        "import oracle_ctypes; from pyjurassic import jtypes as J\n"
        "from test_oracle_geo_atm import ctl_ip, column_atm, at\n"
        "lib = oracle_ctypes.load(); ctl = ctl_ip(lib, 2)\n"
        "atm = column_atm(ctl, [0.0, 1.0], [0.0, 0.0], [0.0, 0.0], nz=1)\n"
        "at(lib, ctl, atm, 0.0, 0.5, 0.0)\n" % (sys.path[:4],))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Cannot identify profiles" in r.stdout
