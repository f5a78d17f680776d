"""2-D and 3-D atmospheres (reference src/atmosphere.c:269-483, IP = 2 and 3)
through the GPU forward model and Jacobian, against the oracle on the same
inputs.

The atmosphere is the clear example's profile repeated at latitudes 0, 9, 18
and 27 degrees with temperature offsets of 0, 4, 8 and 12 K (and an O3 scale),
the example's limb rays (observer at 27 N, tangent points near the equator)
cross the first two within the atmosphere: IP 2 mixes the two nearest
profiles, IP 3 averages within CZ / CX.
The FP64 path (option fp64) agrees with the oracle to ~1e-12; the default
FP32 path to the bar of the other clear-air options (1e-4).
"""
import ctypes as C
import subprocess
import sys

import numpy as np
import pytest

import oracle_ctypes
from pyjurassic import jtypes as J
from pyjurassic import lib as L

pytestmark = pytest.mark.gpu

LATS = (0.0, 9.0, 18.0, 27.0)
CASES = {"ip2": ("IP", 2), "ip3": ("IP", 3, "CZ", 2.5, "CX", 1200)}


@pytest.fixture(scope="module")
def glib():
    return L.load()


@pytest.fixture(scope="module")
def olib():
    return oracle_ctypes.load()


def transect(lib, prefix, keys):
    """ctl, obs and the four-profile atmosphere read with `lib`'s readers."""
    ctl, obs, base = J.ctl_t(), J.obs_t(), J.atm_t()
    argc, argv = J.argv("clear-air.ctl", *keys)
    getattr(lib, prefix + "read_ctl")(argc, argv, C.byref(ctl))
    getattr(lib, prefix + "read_obs")(None, b"org/obs.tab", C.byref(ctl), C.byref(obs))
    getattr(lib, prefix + "read_atm")(None, b"atm_pwin_0_80.tab", C.byref(ctl), C.byref(base))
    atm = J.atm_t()
    n = 0
    for i, lat in enumerate(LATS):
        for l in range(base.np):
            atm.time[n], atm.z[n], atm.lon[n], atm.lat[n] = base.time[l], base.z[l], 0.0, lat
            atm.p[n], atm.t[n] = base.p[l], base.t[l] + 4.0 * i
            for g in range(ctl.ng):
                atm.q[g][n] = base.q[g][l] * (1 + 0.1 * i if g == 2 else 1)
            n += 1
    atm.np = n
    return ctl, obs, atm


def forward(lib, prefix, keys):
    ctl, obs, atm = transect(lib, prefix, keys)
    getattr(lib, prefix + "formod")(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(J.aero_t()))
    rad = np.array([obs.rad[0][i] for i in range(obs.nr)])
    tau = np.array([obs.tau[0][i] for i in range(obs.nr)])
    return rad, tau


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("fp64", [0, 1])
def test_formod_vs_oracle(glib, olib, in_examples, name, fp64):
    print("case %s fp64 %d" % (name, fp64), flush=True)
    keys = CASES[name]
    glib.jur_set_option(b"fp64", fp64)
    try:
        g = forward(glib, "", keys)
    finally:
        glib.jur_set_option(b"fp64", 0)
    o = forward(olib, "oracle_", keys)
    rel = np.max(np.abs(g[0] / o[0] - 1))
    print("%s fp64=%d: max rel rad %.2e, max |dtau| %.2e" % (name, fp64, rel, np.max(np.abs(g[1] - o[1]))),
          flush=True)
    assert rel <= (1e-10 if fp64 else 1e-4)
    np.testing.assert_allclose(g[1], o[1], atol=1e-10 if fp64 else 1e-5)
    # the horizontal structure matters: IP 1 on the first profile differs
    ctl, obs, atm = transect(olib, "oracle_", ())
    atm.np = atm.np // len(LATS)
    olib.oracle_formod(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(J.aero_t()))
    r1 = np.array([obs.rad[0][i] for i in range(obs.nr)])
    assert np.max(np.abs(r1 / o[0] - 1)) > 1e-3


def jacobian(lib, prefix, keys):
    ctl, obs, atm = transect(lib, prefix, keys + ("RETT_ZMIN", 0, "RETT_ZMAX", 40))
    r, aero = C.byref, J.aero_t()
    if prefix:
        n = lib.oracle_atm2x(r(ctl), r(atm), r(aero), None, None, None)
        m = lib.oracle_obs2y(r(ctl), r(obs), None, None, None)
        k = np.zeros((m, n))
        lib.oracle_kernel(r(ctl), r(atm), r(obs), r(aero), k.ctypes.data_as(C.POINTER(C.c_double)), m, n, n)
    else:
        n = lib.jur_state_size(r(ctl), r(atm), r(aero))
        m = lib.jur_meas_size(r(ctl), r(obs))
        k = np.zeros((m, n))
        lib.jur_kernel(r(ctl), r(atm), r(obs), r(aero), k.ctypes.data_as(C.POINTER(C.c_double)), m, n, n)
    return k


@pytest.mark.parametrize("name", sorted(CASES))
def test_kernel_vs_oracle(glib, olib, in_examples, name):
    """Temperature columns of all four profiles (164 elements)."""
    kg = jacobian(glib, "", CASES[name])
    ko = jacobian(olib, "oracle_", CASES[name])
    assert kg.shape == ko.shape and kg.shape[1] == 4 * 41
    scale = np.max(np.abs(ko), axis=1, keepdims=True)
    err = np.max(np.abs(kg - ko) / scale)
    print("%s kernel %s: max |dK| / row max %.2e" % (name, kg.shape, err), flush=True)
    assert err < 1e-6
    # more than one profile's temperature reaches the radiances (the limb
    # rays leave the atmosphere before the equatorial profile)
    hit = [bool(np.any(ko[:, 41 * i:41 * (i + 1)] != 0)) for i in range(4)]
    print("%s: profiles with non-zero columns %s" % (name, hit), flush=True)
    assert sum(hit) >= 2


@pytest.mark.parametrize("keys,msg", [(("IP", 4), "Unknown interpolation method"),
                                      (("IP", 2, "ONEPOINT", 1), "Cannot identify profiles")])
def test_errors(in_examples, keys, msg):
    code = (
        "import ctypes as C, sys; sys.path[:0] = %r\n"
        "import torch\n"
        "from pyjurassic import lib as L, jtypes as J\n"
        "from test_gpu_geo_atm import transect\n"
        "lib = L.load(); keys = %r\n"
        "ctl, obs, atm = transect(lib, '', tuple(k for k in keys if k not in ('ONEPOINT', 1)))\n"
        "if 'ONEPOINT' in keys: atm.lat[1] = 5.0\n"
        "lib.formod(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(J.aero_t()))\n" % (sys.path[:4], keys))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and msg in r.stdout, r.stdout[-500:]


def scatter_case(lib, prefix, tmp_path, ip):
    """One steep ray from 30 km through the aerosol layer of the aerosol
    workload (two channels, one emitter), single scattering with the direct
    fans of srcfunc_sca_3d, two profiles 5 degrees apart (the second 6 K
    warmer); coarse ray steps keep the oracle's 32,400 pencils per point
    affordable.  For IP 3, CX covers every point a direct pencil reaches
    (limb pencils run ~1500 km): without a grid point in range the
    reference's average is 0 / 0 and the NaN ray ends in "Too many LOS
    points!"."""
    from pyjurassic import synth
    tmp_path.mkdir(parents=True, exist_ok=True)
    w = synth.make("aerosol_limb_8192", tmp_path, nray=2).select(nd=2, ng=1)
    w.write_ctl(tmp_path / "c.ctl")
    keys = ["IP", ip, "RAYDZ", 2, "RAYDS", 20] + (["CZ", 3, "CX", 5000] if ip == 3 else [])
    ctl = J.ctl_t()
    argc, argv = J.argv(tmp_path / "c.ctl", *keys)
    getattr(lib, prefix + "read_ctl")(argc, argv, C.byref(ctl))
    aero = J.aero_t()
    getattr(lib, prefix + "read_aero")(None, w.aerofile.encode(), C.byref(ctl), C.byref(aero))
    getattr(lib, prefix + "get_opt_prop")(C.byref(ctl), C.byref(aero))
    atm = J.atm_t()
    n = 0
    for lat, dt in ((0.0, 0.0), (5.0, 6.0)):
        for l in range(w.nlev):
            atm.z[n], atm.lon[n], atm.lat[n], atm.p[n], atm.t[n] = w.z[0, l], 0.0, lat, w.p[0, l], w.t[0, l] + dt
            atm.q[0][n] = w.q[0, 0, l]
            n += 1
    atm.np = n
    obs = J.obs_t()
    obs.nr = 1
    obs.obsz[0], obs.obslon[0], obs.obslat[0] = 30.0, 0.0, 2.0
    obs.vpz[0], obs.vplon[0], obs.vplat[0] = 0.0, 0.0, 2.3
    getattr(lib, prefix + "formod")(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(aero))
    return np.array([obs.rad[i][0] for i in range(2)]), np.array([obs.tau[i][0] for i in range(2)])


@pytest.mark.parametrize("ip", [2, 3])
def test_scattering_3d_vs_oracle(glib, olib, tmp_path, ip):
    """srcfunc_sca_3d (scatter.c:811-895) on the GPU: every scattering point's
    32,400 direct pencils traced and integrated as one level of the
    scattering tree, summed by k_angular3d."""
    print("case scattering ip %d" % ip, flush=True)
    g = scatter_case(glib, "", tmp_path / "g", ip)
    pencils = glib.jur_stat(b"pencils")
    o = scatter_case(olib, "oracle_", tmp_path / "o", ip)
    rel = np.max(np.abs(g[0] / o[0] - 1))
    print("ip %d: %d pencils, rad gpu %s oracle %s, max rel %.2e" % (ip, pencils, g[0], o[0], rel), flush=True)
    assert pencils > 32400  # the primary, its points' direct fans (and solar pencils)
    assert rel < 1e-4
    np.testing.assert_allclose(g[1], o[1], atol=1e-5)


@pytest.mark.parametrize("name", sorted(CASES))
def test_raytrace_los_vs_oracle(glib, olib, in_examples, name):
    """raytrace() on the transect: the LOS points, their atmosphere (IP 2 / 3
    interpolation at each point's longitude and latitude) and column
    densities against the oracle."""
    ctl, obs, atm = transect(olib, "oracle_", CASES[name])
    aero = J.aero_t()
    lo, lg = J.los_t(), J.los_t()
    olib.oracle_raytrace(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(aero), C.byref(lo), 4)
    glib.raytrace(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(aero), C.byref(lg), 4)
    n = lo.np
    assert lg.np == n and n > 50
    for f in ["z", "lon", "lat", "p", "t", "ds"]:
        a = np.frombuffer(getattr(lo, f), dtype=np.float64)[:n]
        b = np.frombuffer(getattr(lg, f), dtype=np.float64)[:n]
        np.testing.assert_allclose(b, a, rtol=1e-11, atol=1e-10, err_msg=f)
    uo = np.array([[lo.u[i][g] for g in range(ctl.ng)] for i in range(n)])
    ug = np.array([[lg.u[i][g] for g in range(ctl.ng)] for i in range(n)])
    np.testing.assert_allclose(ug, uo, rtol=1e-11)
    lat = np.frombuffer(lo.lat, dtype=np.float64)[:n]
    assert lat.max() - lat.min() > 3  # the LOS crosses profiles


@pytest.mark.parametrize("name", sorted(CASES))
def test_hydrostatic_profiles_vs_oracle(glib, olib, in_examples, name):
    """HYDZ rebuilds the pressure of each profile (runs of equal lon, lat)
    from its own reference level; FP64 formod on the rebalanced transect."""
    keys = CASES[name] + ("HYDZ", 12)
    glib.jur_set_option(b"fp64", 1)
    try:
        ctl, obs, atm = transect(glib, "", keys)
        glib.formod(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(J.aero_t()))
        g = np.array([obs.rad[0][i] for i in range(obs.nr)]), np.array(atm.p[:atm.np])
    finally:
        glib.jur_set_option(b"fp64", 0)
    ctl, obs, atm = transect(olib, "oracle_", keys)
    olib.oracle_formod(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(J.aero_t()))
    o = np.array([obs.rad[0][i] for i in range(obs.nr)]), np.array(atm.p[:atm.np])
    np.testing.assert_allclose(g[1], o[1], rtol=1e-12)
    rel = np.max(np.abs(g[0] / o[0] - 1))
    print("%s HYDZ fp64: max rel rad %.2e" % (name, rel), flush=True)
    assert rel <= 1e-10
    # the four profiles were rebalanced separately: their pressures differ
    nl = atm.np // len(LATS)
    assert not np.allclose(o[1][:nl], o[1][3 * nl:], rtol=1e-6)
