"""Pins of the oracle's configuration-parameter grids (no reference
counterpart: the reference fixes 28 fan pencils, 170 Gauss-Hermite radius
nodes and 181 phase angles, src/scatter.c:680, 698-715, jurassic.h:190-193).

* Fan: for 28 pencils the oracle's zenith angles are the reference's
  (scatter.c:698-715, listed here from that loop); the 32-pencil grid (the
  cirrus configuration) is increasing in [0, 180] deg, keeps the
  reference's horizon band and first steps, and its tail steps are equal;
  the isothermal opaque closure (J = B(T)) holds with it.
* Mie at 180 angles (the aerosol configuration): Q_ext (from the forward
  amplitude) and Q_sca equal the reference bhmie's; the Rayleigh limit
  3/4 (1 + cos^2 theta); the phase function normalised over the sphere; for
  smooth phase functions the 180 cell values agree with the reference
  bhmie's 181 nodes interpolated to the cell centres, and the resampled
  nodes with the reference's nodes.
* The 64-radius trapezoid rule integrates the same log-normal mode as the
  reference's Gauss-Hermite rule: the layer optics of the aerosol
  configuration agree between the two grids.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_ctypes
from helpers import read_workload
from pyjurassic import jtypes as J
from pyjurassic import synth


@pytest.fixture(scope="module")
def olib():
    return oracle_ctypes.load()


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def alpha(olib, n):
    a = np.zeros(n)
    olib.oracle_alpha_grid(n, _dp(a))
    return np.degrees(a)


def test_reference_fan_28(olib):
    # scatter.c:698-715: 83..90 in 1 deg steps, then steps 3, 3, 5, 8, 13, 13, ...
    ref = [0, 10, 23, 36, 49, 62, 70, 75, 78, 81] + list(range(83, 91)) + [92, 95, 98, 103, 111, 124, 137, 150,
                                                                             163, 180]
    assert len(ref) == 28
    np.testing.assert_allclose(alpha(olib, 28), ref, atol=1e-12)


@pytest.mark.parametrize("n", [30, 32])
def test_project_fan_grid(olib, n):
    a = alpha(olib, n)
    assert a[0] == 0 and abs(a[-1] - 180) < 1e-12
    assert np.all(np.diff(a) > 0)
    np.testing.assert_allclose(a[n // 2 - 4: n // 2 + 4], np.arange(83, 91), atol=1e-12)
    np.testing.assert_allclose(a[n // 2 + 4: n // 2 + 9], [92, 95, 98, 103, 111], atol=1e-12)
    np.testing.assert_allclose(a[n // 2 - 9: n // 2 - 4], [62, 70, 75, 78, 81], atol=1e-12)
    up, down = np.diff(a[n // 2 + 8:]), np.diff(a[: n // 2 - 8])
    np.testing.assert_allclose(up, up[0], rtol=1e-12)
    np.testing.assert_allclose(down, down[0], rtol=1e-12)
    assert up.max() < 13 and down.max() < 13  # finer than the reference's 13-17 deg tail


def bh(olib, x, m):
    ph181, qe, qs = np.zeros(181), C.c_double(), C.c_double()
    olib.oracle_bhmie(x, m.real, m.imag, _dp(ph181), C.byref(qe), C.byref(qs))
    p180, r181, qe2, qs2 = np.zeros(180), np.zeros(181), C.c_double(), C.c_double()
    olib.oracle_bhmie180(x, m.real, m.imag, _dp(p180), _dp(r181), C.byref(qe2), C.byref(qs2))
    return ph181, qe.value, qs.value, p180, r181, qe2.value, qs2.value


@pytest.mark.parametrize("m", [1.5 + 0.01j, 1.33 + 0.3j, 1.2 + 0.0j])
@pytest.mark.parametrize("x", [0.1, 1.0, 5.0, 20.0, 100.0, 500.0])
def test_bhmie180_efficiencies_equal_reference(olib, x, m):
    _, qe, qs, _, _, qe2, qs2 = bh(olib, x, m)
    assert abs(qe2 / qe - 1) < 1e-11 and abs(qs2 / qs - 1) < 1e-12


def test_bhmie180_rayleigh_limit(olib):
    x, m = 1e-3, 1.33 + 0.0j
    _, _, _, p180, _, qe2, qs2 = bh(olib, x, m)
    th = np.radians(np.arange(180) + 0.5)
    np.testing.assert_allclose(p180, 0.75 * (1 + np.cos(th) ** 2), rtol=1e-5)
    k = (m ** 2 - 1) / (m ** 2 + 2)
    np.testing.assert_allclose(qs2, 8 / 3 * x ** 4 * abs(k) ** 2, rtol=1e-4)


@pytest.mark.parametrize("x", [0.3, 1.0, 3.0, 8.0])
def test_bhmie180_normalised_and_on_reference_curve(olib, x):
    ph181, _, _, p180, r181, _, _ = bh(olib, x, 1.45 + 0.05j)
    th = np.radians(np.arange(180) + 0.5)
    # (1 / 4 pi) int P dOmega = 1/2 int P sin theta dtheta (midpoint rule,
    # whose error grows with the forward peak)
    assert abs(0.5 * np.sum(p180 * np.sin(th)) * np.radians(1.0) - 1) < (2e-4 if x <= 3 else 1e-3)
    if x <= 3:  # smooth enough for linear interpolation over 1 deg to hold to 2e-3
        ref_cells = np.interp(np.arange(180) + 0.5, np.arange(181), ph181)
        np.testing.assert_allclose(p180, ref_cells, rtol=2e-3)
        np.testing.assert_allclose(r181[1:-1], ph181[1:-1], rtol=2e-3)


def test_mie_grid_mode_integral(olib, tmp_path):
    """The aerosol configuration's layer (median radius U(0.1, 1) um, width
    U(1.5, 2)) on its 32 channels: 64-radius trapezoid rule and 180 angles
    against the reference's 170 Gauss-Hermite nodes and 181 angles."""
    w = synth.make("aerosol_limb_8192", tmp_path, nray=2)
    out = {}
    for grid in (0, 1):
        olib.oracle_set_mie_grid(grid)
        try:
            ctl, aero = read_workload(olib, "oracle_", w, tmp_path / ("c%d.ctl" % grid))
        finally:
            olib.oracle_set_mie_grid(0)
        nd = ctl.nd
        out[grid] = (np.array([aero.beta_e[0][i] for i in range(nd)]),
                     np.array([aero.beta_s[0][i] for i in range(nd)]),
                     np.array([[aero.p[0][i][t] for t in range(J.NTHETA)] for i in range(nd)]))
    for a, b in zip(out[1][:2], out[0][:2]):
        np.testing.assert_allclose(a, b, rtol=1e-3)
    np.testing.assert_allclose(out[1][2][:, 1:-1], out[0][2][:, 1:-1], rtol=5e-3)
