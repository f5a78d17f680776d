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
