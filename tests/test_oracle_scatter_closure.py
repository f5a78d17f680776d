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
with 28 pencils.
"""
import ctypes as C
import struct

import numpy as np
import pytest

import oracle_ctypes
from helpers import read_workload
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
