"""Closed-form pins of the two oracle functions no reference fixture exercises:
the field-of-view convolution (formod_fov, reference src/forwardmodel.c:156-220)
and the solar single-scattering term by day (srcfunc_sca_sun,
src/scatter.c:899-995; every reference fixture is at t = 0, night at
longitude 0).

FOV: radiances and transmittances are set directly in obs_t and convolved
by the oracle alone (oracle_formod_fov).
* A one-node (delta) shape leaves every value unchanged.
* Values linear in tangent height are linear in the offset, so any shape
  gives the value at the weight-averaged offset sum(w dz) / sum(w), also
  where the offsets reach beyond the profile (linear extrapolation); the
  weights here do not sum to one, so a dropped normalisation fails.
* Values that are not linear: the weighted mean of numpy's linear
  interpolation (np.interp) over the rays of the same time within +-NFOV = 50
  rays; interleaved time groups with different profiles fail if the groups
  were mixed, a shape reaching beyond the +-50-ray window fails if the
  window were dropped.

Sun: a transparent atmosphere (no emitters, layer optics zero) and a phase
function linear in the scattering angle, P(theta) = 1 + 2 theta / pi, which
the 181-node linear interpolation reproduces exactly.  The direction to the
sun is built here from suncoord's zenith and azimuth angles (pinned by the
equinox / solstice checks of test_oracle_math.py) in a textbook east-north-up
frame; the LOS tangent dx is placed at a known angle theta0 from it, so the
term must be J = 6.764e-5 P(theta0) B(5780 K, nu) with B the Planck function.
The sun is at infinity, so its direction is one Earth-fixed vector: the same
dx at a second point, 30 deg away, must give the same J (an azimuth measured
the wrong way round or a mirrored scattering angle fails).  A sun ray that
meets the ground gives 0; one that grazes past the horizon (zenith angle
between 90 deg and 90 deg + the horizon dip) gives the full term.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_ctypes
from helpers import read_workload
from pyjurassic import jtypes as J
from pyjurassic import synth

RE = 6367.421
C1, C2 = 1.19104259e-8, 1.43877506  # reference src/jurassic.h:114-117


@pytest.fixture(scope="module")
def olib():
    return oracle_ctypes.load()


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------------------
# formod_fov
# ---------------------------------------------------------------------------

def _fov_case(tmp_path, shape, tpz, time, rad, tau):
    """ctl with the FOV shape file, obs with the given tangent heights,
    times and values (rad, tau [nd][nr])."""
    fov = tmp_path / "fov.tab"
    fov.write_text("# dz [km] weight\n" + "".join("%.17g %.17g\n" % (d, w) for d, w in shape))
    ctl = J.ctl_t()
    ctl.nd = rad.shape[0]
    ctl.fov = str(fov).encode()
    obs = J.obs_t()
    obs.nr = len(tpz)
    for ir in range(obs.nr):
        obs.tpz[ir] = tpz[ir]
        obs.time[ir] = time[ir]
        for id_ in range(ctl.nd):
            obs.rad[id_][ir] = rad[id_, ir]
            obs.tau[id_][ir] = tau[id_, ir]
    return ctl, obs


def _values(obs, nd):
    nr = obs.nr
    rad = np.array([[obs.rad[i][r] for r in range(nr)] for i in range(nd)])
    tau = np.array([[obs.tau[i][r] for r in range(nr)] for i in range(nd)])
    return rad, tau


def test_fov_delta_shape_is_identity(olib, tmp_path):
    rng = np.random.default_rng(1)
    nr, nd = 23, 3
    tpz = np.sort(rng.uniform(5, 40, nr))[::-1].copy()
    rad = rng.uniform(1e-4, 1e-2, (nd, nr))
    tau = rng.uniform(0.1, 1.0, (nd, nr))
    ctl, obs = _fov_case(tmp_path, [(0.0, 2.5)], tpz, np.zeros(nr), rad, tau)
    olib.oracle_formod_fov(C.byref(ctl), C.byref(obs))
    r2, t2 = _values(obs, nd)
    np.testing.assert_allclose(r2, rad, rtol=1e-14, atol=0)
    np.testing.assert_allclose(t2, tau, rtol=1e-14, atol=0)


@pytest.mark.parametrize("decreasing", [False, True])
def test_fov_linear_profile_gives_value_at_mean_offset(olib, tmp_path, decreasing):
    rng = np.random.default_rng(2)
    nr, nd = 30, 2
    tpz = np.sort(rng.uniform(8, 30, nr))
    if decreasing:
        tpz = tpz[::-1].copy()
    a, b = np.array([3e-3, 1e-2]), np.array([-4e-5, 2e-4])
    at, bt = np.array([0.9, 0.2]), np.array([-0.01, 0.02])
    rad = a[:, None] + b[:, None] * tpz[None, :]
    tau = at[:, None] + bt[:, None] * tpz[None, :]
    # asymmetric shape, weights summing to 3.7 (not 1), offsets up to 3 km:
    # the outermost rays extrapolate
    shape = [(-2.0, 0.4), (-0.5, 1.3), (0.0, 1.0), (0.7, 0.6), (3.0, 0.4)]
    dz_mean = sum(d * w for d, w in shape) / sum(w for _, w in shape)
    ctl, obs = _fov_case(tmp_path, shape, tpz, np.zeros(nr), rad, tau)
    olib.oracle_formod_fov(C.byref(ctl), C.byref(obs))
    r2, t2 = _values(obs, nd)
    z = tpz + dz_mean
    np.testing.assert_allclose(r2, a[:, None] + b[:, None] * z[None, :], rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(t2, at[:, None] + bt[:, None] * z[None, :], rtol=1e-12, atol=1e-15)


def _expected_nonlinear(tpz, time, vals, shape):
    """Weighted mean over the shape of the piecewise-linear profile of the
    rays with the same time within +-50 rays (np.interp inside the
    profile's range, linear continuation of the end segments beyond)."""
    nr = len(tpz)
    out = np.zeros_like(vals)
    ws = sum(w for _, w in shape)
    for ir in range(nr):
        sel = [r for r in range(max(ir - 50, 0), min(ir + 51, nr)) if time[r] == time[ir]]
        zs = tpz[sel]
        o = np.argsort(zs)
        zs = zs[o]
        for k in range(vals.shape[0]):
            vs = vals[k, sel][o]
            acc = 0.0
            for d, w in shape:
                x = tpz[ir] + d
                if x < zs[0]:
                    v = vs[0] + (vs[1] - vs[0]) / (zs[1] - zs[0]) * (x - zs[0])
                elif x > zs[-1]:
                    v = vs[-2] + (vs[-1] - vs[-2]) / (zs[-1] - zs[-2]) * (x - zs[-2])
                else:
                    v = np.interp(x, zs, vs)
                acc += w * v
            out[k, ir] = acc / ws
    return out


def test_fov_weighted_mean_within_time_groups(olib, tmp_path):
    """Two interleaved scans (times 0 and 60 s) with different profiles."""
    nr = 40
    time = np.where(np.arange(nr) % 2 == 0, 0.0, 60.0)
    tpz = np.where(time == 0, 35.0 - 0.7 * np.arange(nr), 34.6 - 0.65 * np.arange(nr))
    rad = np.vstack([1e-3 * np.exp(-tpz / 6.0) * (1 + (time > 0)), 2e-3 / (1 + 0.01 * tpz ** 2)])
    tau = np.vstack([np.exp(-30 / (tpz + 5)), 0.5 + 0.4 * np.sin(tpz / 5 + time / 60)])
    shape = [(-1.0, 0.2), (-0.4, 0.9), (0.0, 1.2), (0.3, 0.8), (1.1, 0.1)]
    ctl, obs = _fov_case(tmp_path, shape, tpz, time, rad, tau)
    olib.oracle_formod_fov(C.byref(ctl), C.byref(obs))
    r2, t2 = _values(obs, 2)
    np.testing.assert_allclose(r2, _expected_nonlinear(tpz, time, rad, shape), rtol=1e-12)
    np.testing.assert_allclose(t2, _expected_nonlinear(tpz, time, tau, shape), rtol=1e-12)
    # mixing the two scans would move the result by far more than that
    mixed = _expected_nonlinear(tpz, np.zeros(nr), rad, shape)
    assert np.max(np.abs(mixed / r2 - 1)) > 1e-3


def test_fov_window_of_fifty_rays(olib, tmp_path):
    """200 rays 0.05 km apart: the +-50-ray window spans +-2.5 km, a 4 km
    offset extrapolates from its end segment, not from the far rays."""
    nr = 200
    tpz = 5.0 + 0.05 * np.arange(nr)
    rad = (1e-3 * np.exp(-((tpz - 10.0) / 3.0) ** 2))[None, :]
    tau = (0.5 + 0.5 * np.tanh(tpz - 9.0))[None, :]
    shape = [(-4.0, 0.5), (0.0, 1.0), (4.0, 0.5)]
    ctl, obs = _fov_case(tmp_path, shape, tpz, np.zeros(nr), rad, tau)
    olib.oracle_formod_fov(C.byref(ctl), C.byref(obs))
    r2, t2 = _values(obs, 1)
    np.testing.assert_allclose(r2, _expected_nonlinear(tpz, np.zeros(nr), rad, shape), rtol=1e-12, atol=1e-20)
    np.testing.assert_allclose(t2, _expected_nonlinear(tpz, np.zeros(nr), tau, shape), rtol=1e-12, atol=1e-18)
    # the whole profile (no window) differs at the middle rays
    ws = sum(w for _, w in shape)
    full = np.array([sum(w * np.interp(z + d, tpz, rad[0]) for d, w in shape) / ws for z in tpz])
    assert np.max(np.abs(full[80:120] / r2[0, 80:120] - 1)) > 1e-2


# ---------------------------------------------------------------------------
# srcfunc_sca_sun by day
# ---------------------------------------------------------------------------

def _enu(lon, lat):
    lo, la = np.radians(lon), np.radians(lat)
    up = np.array([np.cos(la) * np.cos(lo), np.cos(la) * np.sin(lo), np.sin(la)])
    east = np.array([-np.sin(lo), np.cos(lo), 0.0])
    north = np.array([-np.sin(la) * np.cos(lo), -np.sin(la) * np.sin(lo), np.cos(la)])
    return east, north, up


def _sun_dir(olib, sec, lon, lat):
    """Unit vector towards the sun (Earth-fixed) from suncoord's azimuth
    (from north, clockwise) and zenith angle at (lon, lat)."""
    azi, sza = C.c_double(), C.c_double()
    olib.oracle_suncoord(sec, lon, lat, C.byref(azi), C.byref(sza))
    e, n, u = _enu(lon, lat)
    a, s = np.radians(azi.value), np.radians(sza.value)
    return np.sin(s) * np.sin(a) * e + np.sin(s) * np.cos(a) * n + np.cos(s) * u, sza.value


def _time_for_sza(olib, lon, lat, target):
    """The time of 2000-01-01 (seconds) whose solar zenith angle at (lon,
    lat) is closest to target, in the morning."""
    best, bsec = 1e9, 0.0
    for sec in np.arange(0.0, 43200.0, 20.0):
        azi, sza = C.c_double(), C.c_double()
        olib.oracle_suncoord(sec, lon, lat, C.byref(azi), C.byref(sza))
        if abs(sza.value - target) < best:
            best, bsec = abs(sza.value - target), sec
    return bsec


def _planck(t, nu):
    return C1 * nu ** 3 / np.expm1(C2 * nu / t)


@pytest.fixture(scope="module")
def sun_setup(olib, tmp_path_factory):
    """The aerosol workload's channels and layer (15-20 km), without
    emitters, with zero layer optics and P(theta) = 1 + 2 theta / pi;
    straight rays."""
    tmp = tmp_path_factory.mktemp("sun")
    w = synth.make("aerosol_limb_8192", tmp, nray=2).select(nd=2, ng=1)
    ctl, aero = read_workload(olib, "oracle_", w, tmp / "c.ctl")
    ctl.ng = 0
    ctl.refrac = 0
    for id_ in range(ctl.nd):
        aero.beta_e[0][id_] = aero.beta_s[0][id_] = aero.beta_a[0][id_] = 0.0
        for it in range(J.NTHETA):
            aero.p[0][id_][it] = 1.0 + it / 90.0  # 1 + 2 theta / pi at theta = it deg
    atm = J.atm_t()
    atm.np = w.nlev
    for l in range(w.nlev):
        atm.z[l], atm.p[l], atm.t[l] = w.z[0, l], w.p[0, l], 100.0
    return ctl, atm, aero


def _sun_term(olib, setup, sec, lon, lat, z, dx):
    ctl, atm, aero = setup
    e, n, u = _enu(lon, lat)
    x = (RE + z) * u
    src = np.zeros(ctl.nd)
    olib.oracle_reset()
    try:
        olib.oracle_srcfunc_sca_sun(C.byref(ctl), C.byref(atm), C.byref(aero), sec, _dp(np.ascontiguousarray(x)),
                                    _dp(np.ascontiguousarray(dx)), 0, _dp(src))
    finally:
        olib.oracle_reset()
    return src


def _at_angle(v, theta0, other):
    """Unit vector at angle theta0 from unit vector v, in the plane of v and other."""
    w = other - np.dot(other, v) * v
    w /= np.linalg.norm(w)
    return np.cos(theta0) * v + np.sin(theta0) * w


def test_sun_azimuth_convention(olib):
    """Near the March equinox the sun rises in the east and sets in the west
    at the equator (azimuth from north, clockwise)."""
    sec = 79.5 * 86400  # 2000-03-20 12:00 UTC + 0 d
    azi, sza = C.c_double(), C.c_double()
    olib.oracle_suncoord(sec - 5 * 3600, 0.0, 0.0, C.byref(azi), C.byref(sza))
    assert abs(azi.value - 90) < 2 and 70 < sza.value < 80
    olib.oracle_suncoord(sec + 5 * 3600, 0.0, 0.0, C.byref(azi), C.byref(sza))
    assert abs(abs(azi.value) - 90) < 2 and azi.value < 0 and 70 < sza.value < 80


@pytest.mark.parametrize("target_sza", [25.0, 70.0])
@pytest.mark.parametrize("theta0_deg", [10.0, 90.0, 155.0])
def test_sun_term_closed_form(olib, sun_setup, target_sza, theta0_deg):
    ctl = sun_setup[0]
    lon, lat = 10.0, 5.0
    sec = _time_for_sza(olib, lon, lat, target_sza)
    ek, sza = _sun_dir(olib, sec, lon, lat)
    assert sza < 85
    th0 = np.radians(theta0_deg)
    dx = _at_angle(ek, th0, np.array([0.3, -0.5, 0.8]))
    nu = np.array([ctl.nu[i] for i in range(ctl.nd)])
    expect = 6.764e-5 * (1 + 2 * th0 / np.pi) * _planck(5780.0, nu)
    got = _sun_term(olib, sun_setup, sec, lon, lat, 17.0, dx)
    np.testing.assert_allclose(got, expect, rtol=1e-9)
    # the same Earth-fixed direction seen from a point 30 deg away
    got2 = _sun_term(olib, sun_setup, sec, lon + 25.0, lat + 17.0, 17.0, dx)
    np.testing.assert_allclose(got2, expect, rtol=1e-9)


def test_sun_below_horizon(olib, sun_setup):
    """At 17 km the horizon dips by acos(RE / (RE + 17)) = 4.18 deg: a sun
    2 deg below the local horizontal is still seen over the limb (full
    term); 10 deg below, its ray meets the ground (no term)."""
    ctl = sun_setup[0]
    lon, lat = 10.0, 5.0
    nu = np.array([ctl.nu[i] for i in range(ctl.nd)])
    dip = np.degrees(np.arccos(RE / (RE + 17.0)))
    assert 4.1 < dip < 4.3
    for target, seen in ((92.0, True), (100.0, False)):
        sec = _time_for_sza(olib, lon, lat, target)
        ek, sza = _sun_dir(olib, sec, lon, lat)
        assert abs(sza - target) < 0.2
        dx = _at_angle(ek, np.radians(40.0), np.array([0.0, 0.0, 1.0]))
        got = _sun_term(olib, sun_setup, sec, lon, lat, 17.0, dx)
        if seen:
            np.testing.assert_allclose(got, 6.764e-5 * (1 + 80.0 / 180) * _planck(5780.0, nu), rtol=1e-9)
        else:
            assert np.all(got == 0.0)
