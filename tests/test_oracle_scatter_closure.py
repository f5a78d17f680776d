"""Closure of the scattering source in an isothermal, nearly opaque atmosphere.

In an isothermal atmosphere at T whose lines of sight all become opaque, every
secondary pencil sees the Planck radiance B(T) (up to its residual
transmittance); the angular average J (reference src/scatter.c:753-803,
normalised by the sum of its weights) must then equal B(T), the scattering
source ((beta - af beta_s) B + af beta_s J) / beta equals B, and the
radiance of a pencil telescopes to B(T) (1 - tau) whatever the emissivities
of its segments.  The observers sit inside the aerosol (or cloud) layer, so
scattering points come first on their paths where they weigh most.  A wrong
weight normalisation, a dropped sin(theta) or a misindexed fan pencil breaks
the equality; with a J of 2 B the radiance would be off by more than 1e-4
here.  This pins srcfunc_sca_1d at scattering depth 1 and 2 and for the fan
sizes the oracle supports, which no reference fixture covers beyond depth 1
with 28 pencils, and srcfunc_sca_3d (IP 2 / 3) with its normalisation over
all channels.
"""
import ctypes as C
import struct

import numpy as np
import pytest

import oracle_ctypes
from helpers import read_workload
from pyjurassic import jtypes as J
from pyjurassic import synth

T_ISO = 250.0


@pytest.fixture(scope="module")
def olib():
    return oracle_ctypes.load()


def write_opaque_table(path):
    """A table whose emissivity approaches 1 for the column densities of LOS
    segments but stays below it (reference binary layout,
    src/forwardmodel.c:636-651): a segment of zero gas transmittance adds no
    radiance in the reference (forwardmodel.c:366), so the path must reach
    opacity over a few segments, not in one."""
    p = np.array([1e-6, 1.0, 2000.0])
    t = np.array([150.0, 250.0, 350.0])
    u = np.logspace(12, 26, 8).astype(np.float32)
    eps = np.array([0.5, 0.9, 0.99, 0.999, 0.9999, 0.99999, 0.999999, 0.9999999], dtype=np.float32)
    parts = [struct.pack("<i", len(p)), p.astype("<f8").tobytes(), np.full(len(p), len(t), "<i4").tobytes()]
    for _ in p:
        parts.append(t.astype("<f8").tobytes())
        parts.append(np.full(len(t), len(u), "<i4").tobytes())
        for _ in t:
            parts.append(u.astype("<f4").tobytes())
            parts.append(eps.astype("<f4").tobytes())
    path.write_bytes(b"".join(parts))


@pytest.mark.parametrize("workload,zlayer,depth,fan", [("aerosol_limb_8192", 17.0, 1, 28),
                                                       ("aerosol_limb_8192", 17.0, 1, 16),
                                                       ("cirrus_4096", 10.5, 1, 32),
                                                       ("cirrus_4096", 10.5, 2, 10)])
def test_isothermal_opaque_radiance_is_planck(olib, tmp_path, workload, zlayer, depth, fan):
    w = synth.make(workload, tmp_path, nray=2)
    w = w.select(nd=2, ng=1)
    for f in tmp_path.glob("*.bin"):
        write_opaque_table(f)
    w.t[:] = T_ISO
    # observers inside the layer at midnight (no solar term): one view
    # steeply down, one nearly horizontal
    g = np.zeros((7, 2))
    g[1] = zlayer
    g[4] = [0.0, zlayer - 0.1]
    g[6] = [0.5, 3.0]
    w.geom = g
    ctl, aero = read_workload(olib, "oracle_", w, tmp_path / "c.ctl")
    olib.oracle_reset()
    olib.oracle_set_fan(fan)
    try:
        n, nd = w.nray, ctl.nd
        rad, tau = np.zeros((n, nd)), np.zeros((n, nd))
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        z, p, t, q = f(w.z), f(w.p), f(w.t), f(w.q)
        gg = [f(w.geom[i]) for i in range(7)]
        olib.oracle_formod_batch(C.byref(ctl), n, w.nlev, dp(z), dp(p), dp(t), dp(q), None, *[dp(x) for x in gg],
                                 0, n, 0, dp(rad), dp(tau), None, None, None, None, C.byref(aero), depth)
        src = np.zeros(nd)
        olib.oracle_srcfunc_planck(C.byref(ctl), T_ISO, dp(src))
    finally:
        olib.oracle_set_fan(28)
        olib.oracle_reset()
    # every pencil, the fan pencils included, ends nearly opaque; J then
    # differs from B(T) by their residual transmittance (<= 1e-4).  Ray 0
    # ends on the ground, whose emission B(T) tau completes the sum to B(T);
    # ray 1 leaves to space with B(T) (1 - tau).
    assert np.all(tau < 1e-4)
    expect = src[None, :] * np.array([[1.0], [0.0]]) + src[None, :] * (1 - tau) * np.array([[0.0], [1.0]])
    dev = np.max(np.abs(rad / expect - 1))
    print("max deviation from the closure %.2e" % dev)
    assert dev < 1e-7


def _geo_setup(olib, tmp_path, nu1=None):
    """The aerosol workload reduced to one emitter, opaque tables, isothermal,
    as two identical profiles 5 degrees apart (IP 2)."""
    w = synth.make("aerosol_limb_8192", tmp_path, nray=2)
    w = w.select(nd=2, ng=1)
    for f in tmp_path.glob("*.bin"):
        write_opaque_table(f)
    w.write_ctl(tmp_path / "c.ctl")
    keys = ["IP", 2] + (["NU[1]", "%.4f" % nu1(w)] if nu1 else [])
    ctl = J.ctl_t()
    argc, argv = J.argv(tmp_path / "c.ctl", *keys)
    olib.oracle_read_ctl(argc, argv, C.byref(ctl))
    aero = J.aero_t()
    olib.oracle_read_aero(None, w.aerofile.encode(), C.byref(ctl), C.byref(aero))
    olib.oracle_get_opt_prop(C.byref(ctl), C.byref(aero))
    atm = J.atm_t()
    n = 0
    for lat in (0.0, 5.0):
        for l in range(w.nlev):
            atm.z[n], atm.lon[n], atm.lat[n], atm.p[n], atm.t[n] = w.z[0, l], 0.0, lat, w.p[0, l], T_ISO
            atm.q[0][n] = w.q[0, 0, l]
            n += 1
    atm.np = n
    return ctl, atm, aero


@pytest.mark.parametrize("same_channels", [False, True])
def test_direct_fan_source(olib, tmp_path, same_channels):
    """srcfunc_sca_3d (IP 2, scatter.c:811-895) inside the aerosol layer of an
    isothermal opaque atmosphere: every one of the 32,400 direct pencils sees
    B(T), so with one channel J = B(T).  The reference normalises by the
    weights summed over all channels (scatter.c:883, 891): with two channels
    of the same wavenumber (identical tables and phase functions) each gets
    exactly J = B(T) / 2."""
    ctl, atm, aero = _geo_setup(olib, tmp_path, (lambda w: w.nu[0]) if same_channels else None)
    if not same_channels:
        ctl.nd = 1
    x = (C.c_double * 3)(6367.421 + 17.0, 0.0, 0.0)
    dx = (C.c_double * 3)(0.0, 1.0, 0.0)
    src, b = np.zeros(ctl.nd), np.zeros(ctl.nd)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    olib.oracle_reset()
    try:
        olib.oracle_srcfunc_sca(C.byref(ctl), C.byref(atm), C.byref(aero), 0.0, x, dx, 0, dp(src), 1)
        olib.oracle_srcfunc_planck(C.byref(ctl), T_ISO, dp(b))
    finally:
        olib.oracle_reset()
    ratio = src / b
    print("J / B(T):", ratio)
    if same_channels:
        assert src[0] == src[1]
        np.testing.assert_allclose(ratio, 0.5, atol=1e-4)
    else:
        assert ctl.nd == 1 and 1 - 1e-4 < ratio[0] <= 1
