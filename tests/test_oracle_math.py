"""The oracle's functions that no reference fixture pins, checked against
what the mathematics fixes (closed forms, limits, invariants)."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_ctypes
from pyjurassic import jtypes as J


@pytest.fixture(scope="module")
def olib():
    return oracle_ctypes.load()


def test_brightness_inverts_planck(olib):
    # brightness(planck(T, nu), nu) == T (forwardmodel.c:7-11, 569-573)
    for nu in [650.0, 792.0, 1200.0, 2300.0]:
        for t in [150.0, 220.0, 288.15, 320.0]:
            r = olib.oracle_planck(t, nu)
            assert olib.oracle_brightness(r, nu) == pytest.approx(t, rel=1e-13)
    # planck against the textbook form with the same constants
    nu, t = 792.0, 250.0
    assert olib.oracle_planck(t, nu) == pytest.approx(1.19104259e-8 * nu ** 3 / math.expm1(1.43877506 * nu / t), rel=1e-14)


def test_gauher_moments(olib):
    # Gauss-Hermite with 170 nodes integrates e^{-z^2} z^{2k} exactly:
    # sum w z^{2k} = Gamma(k + 1/2)
    x = (C.c_double * 170)()
    w = (C.c_double * 170)()
    olib.oracle_gauher(x, w)
    xs, ws = np.array(x), np.array(w)
    for k in range(6):
        assert np.sum(ws * xs ** (2 * k)) == pytest.approx(math.gamma(k + 0.5), rel=1e-10)
    assert np.sum(ws * xs ** 3) == pytest.approx(0.0, abs=1e-10)
    assert np.all(np.diff(xs[:85]) < 0)  # nodes descend from the largest


def bhmie(olib, x, nr, ni):
    ph = (C.c_double * 181)()
    qe, qs = C.c_double(), C.c_double()
    olib.oracle_bhmie(x, nr, ni, ph, C.byref(qe), C.byref(qs))
    return np.array(ph), qe.value, qs.value


def test_bhmie_energy_conservation(olib):
    # no absorption: Q_ext == Q_sca
    for x in [0.3, 2.0, 15.0, 80.0]:
        _, qe, qs = bhmie(olib, x, 1.33, 0.0)
        assert qe == pytest.approx(qs, rel=1e-9)


def test_bhmie_phase_normalised(olib):
    # (1/2) int_0^pi P(theta) sin(theta) dtheta == 1 on the 1-degree grid
    th = np.radians(np.arange(181.0))
    for x, nr, ni in [(0.5, 1.5, 0.01), (3.0, 1.33, 0.0), (10.0, 1.4, 0.1)]:
        ph, _, _ = bhmie(olib, x, nr, ni)
        integ = 0.5 * np.trapezoid(ph * np.sin(th), th)
        assert integ == pytest.approx(1.0, rel=2e-3)


def test_bhmie_rayleigh_limit(olib):
    # small spheres: Q_sca = 8/3 x^4 |(m^2-1)/(m^2+2)|^2, P = 3/4 (1 + cos^2)
    x, m = 0.01, complex(1.5, 0.0)
    ph, qe, qs = bhmie(olib, x, m.real, m.imag)
    k = abs((m * m - 1) / (m * m + 2)) ** 2
    assert qs == pytest.approx(8.0 / 3.0 * x ** 4 * k, rel=1e-3)
    th = np.radians(np.arange(181.0))
    np.testing.assert_allclose(ph, 0.75 * (1 + np.cos(th) ** 2), rtol=1e-3)


def test_bhmie_large_sphere_extinction_paradox(olib):
    # Q_ext -> 2 for large absorbing spheres
    _, qe, _ = bhmie(olib, 2000.0, 1.5, 0.5)
    assert qe == pytest.approx(2.0, abs=0.02)


def test_suncoord_equinox_noon(olib):
    # 2000-03-20T12:00Z: the Sun is near the zenith at (lon ~ 0, lat ~ 0)
    sec = (31 + 29 + 19) * 86400.0 + 12 * 3600.0
    azi, sza = C.c_double(), C.c_double()
    olib.oracle_suncoord(sec, 0.0, 0.0, C.byref(azi), C.byref(sza))
    assert sza.value < 3.0
    # midnight at lon 0 on 2000-01-01: Sun below the horizon
    olib.oracle_suncoord(0.0, 0.0, 0.0, C.byref(azi), C.byref(sza))
    assert sza.value > 150.0
    # June solstice noon at the Tropic of Cancer
    sec = (31 + 29 + 31 + 30 + 31 + 20) * 86400.0 + 12 * 3600.0
    olib.oracle_suncoord(sec, 0.0, 23.44, C.byref(azi), C.byref(sza))
    assert sza.value < 2.0


def test_hydrostatic_isothermal_dry(olib):
    # dry isothermal column: p(z) = p(z_ref) exp(-M g(z) (z - z_ref)/(R T)),
    # with the reference's z-dependent gravity averaged over each layer
    ctl, atm = J.ctl_t(), J.atm_t()
    ctl.ng = 0
    ctl.nw = 0
    ctl.hydz = 0.0
    atm.np = 21
    for i in range(21):
        atm.z[i] = float(i)
        atm.t[i] = 250.0
        atm.p[i] = 1000.0 if i == 0 else 1.0
    olib.oracle_hydrostatic(C.byref(ctl), C.byref(atm))
    p = np.array([atm.p[i] for i in range(21)])
    M, R, T = 28.96456e-3, 8.314472, 250.0
    expect = [1000.0]
    for i in range(1, 21):
        zz = np.linspace(i - 1, i, 20)
        g = 9.780318 - 3.086e-3 * zz
        expect.append(expect[-1] * math.exp(-np.mean(M * g / (R * T)) * 1000.0))
    np.testing.assert_allclose(p, expect, rtol=1e-12)


def test_ctm_n2_o2_at_nodes(olib):
    # N2 (continua.c:1980-1992) and O2 (:2037-2048) at T = 296 K, p = P0
    # reduce to 0.1 q b (273/296)^2 (...) with b read at a table node.
    p0 = 1013.25
    t = 296.0
    n2 = olib.oracle_ctmn2(2300.0, p0, t)
    assert n2 == pytest.approx(0.1 * (273 / 296) ** 2 * 0.79 * 1.11e-6 * (0.79 + 0.21 * (1.294 - 0.4545)), rel=1e-12)
    o2 = olib.oracle_ctmo2(1600.0, p0, t)
    assert o2 == pytest.approx(0.1 * (273 / 296) ** 2 * 0.21 * 3.231, rel=1e-12)  # b(1600 cm^-1), table node 48
    assert olib.oracle_ctmn2(792.0, p0, t) == 0.0
    assert olib.oracle_ctmo2(792.0, p0, t) == 0.0
