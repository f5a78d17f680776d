"""The batched entry point (jur_formod_batch) on the synthetic workloads,
against the oracle on the same seeded inputs.

Bars: radiance within 1e-4 relative (<= 0.01 K brightness temperature),
transmittance within 1e-5 absolute, tangent points within 1e-6 km, LOS point
counts equal.  At the full 65,536-ray size the check is batch invariance:
every ray's radiances are bit-identical whether it is computed in the full
batch or in a small one (rays are independent; the sort by path length only
changes which lane computes a ray).
"""
import numpy as np
import pytest

import oracle_ctypes
from helpers import gpu_batch, oracle_batch
from pyjurassic import lib as L
from pyjurassic import synth

pytestmark = pytest.mark.gpu

RAD_RTOL = 1e-4
TAU_ATOL = 1e-5


@pytest.fixture(scope="module")
def glib():
    return L.load()


@pytest.fixture(scope="module")
def olib():
    return oracle_ctypes.load()


def check(got, ref, npts=True):
    rad, tau, tp, n = got
    orad, otau, otp, on = ref
    rel = np.abs(rad / orad - 1.0)
    print("max rel rad %.3e  max abs tau %.3e" % (rel.max(), np.abs(tau - otau).max()), flush=True)
    np.testing.assert_allclose(rad, orad, rtol=RAD_RTOL)
    np.testing.assert_allclose(tau, otau, atol=TAU_ATOL)
    np.testing.assert_allclose(tp, otp, atol=1e-6)
    if npts:
        np.testing.assert_array_equal(n, on)


def test_clear_limb_17(glib, olib, tmp_path):
    """81-level clear-air limb case: 17 tangent heights 6-30 km, CO2 + H2O,
    4 channels 792-832 cm^-1, refraction on."""
    w = synth.make("clear_limb_17", tmp_path)
    print("clear_limb_17: %d rays" % w.nray, flush=True)
    check(gpu_batch(glib, w, tmp_path / "g.ctl"), oracle_batch(olib, w, tmp_path / "o.ctl"))


def test_clear_batch_slice(glib, olib, tmp_path):
    """384 rays of the 65,536-ray limb/nadir batch, 5 emitters, 64 channels."""
    w = synth.make("clear_batch_65536", tmp_path, nray=384)
    print("clear_batch_65536[:384]", flush=True)
    check(gpu_batch(glib, w, tmp_path / "g.ctl"), oracle_batch(olib, w, tmp_path / "o.ctl"))


def test_clear_batch_full_invariance(glib, tmp_path):
    """All 65,536 rays: every ray of a 512-ray subset is bit-identical to the
    same ray computed in the full batch."""
    w = synth.make("clear_batch_65536", tmp_path)
    print("clear_batch_65536 full", flush=True)
    full = gpu_batch(glib, w, tmp_path / "g.ctl")
    sel = np.random.default_rng(7).choice(w.nray, 512, replace=False)
    part = gpu_batch(glib, w.select(rays=sel), tmp_path / "p.ctl")
    for a, b in zip(part, full):
        np.testing.assert_array_equal(a, b[..., sel] if a.ndim == 2 and a.shape[0] == 3 else b[sel])
    assert np.all(np.isfinite(full[0])) and np.all(full[0] > 0)
    assert np.all((full[1] >= 0) & (full[1] <= 1))


def _crossing(w, top, n_limb, n_nadir):
    """Indices of limb rays with tangent height below `top` and nadir rays."""
    limb = np.where((w.geom[4] > 0) & (w.geom[4] < top))[0][:n_limb]
    nadir = np.where((w.geom[4] == 0) & (w.geom[6] != 0))[0][:n_nadir]
    return np.concatenate([limb, nadir])


def test_single_scattering_daytime(glib, olib, tmp_path):
    """Aerosol layer at 15-20 km, single scattering with the solar term (local
    noon), 4 limb rays crossing the layer, 8 channels, 5 emitters."""
    w = synth.make("aerosol_limb_8192", tmp_path, nray=64)
    w = w.select(rays=_crossing(w, 20.0, 4, 0), nd=8)
    print("aerosol_limb: rays", w.nray, flush=True)
    got = gpu_batch(glib, w, tmp_path / "g.ctl")
    print("pencils", glib.jur_stat(b"pencils"), "points", glib.jur_stat(b"scattering_points"), flush=True)
    check(got, oracle_batch(olib, w, tmp_path / "o.ctl"))


def test_depth2_cirrus_daytime(glib, olib, tmp_path):
    """Cirrus at 9-12 km, scattering depth 2 (secondary fans of secondary
    pencils) with the solar term: one limb and one nadir ray, 2 channels,
    CO2 + H2O."""
    w = synth.make("cirrus_4096", tmp_path, nray=64)
    w = w.select(rays=_crossing(w, 12.0, 1, 1), nd=2, ng=2)
    print("cirrus depth 2: rays", w.nray, flush=True)
    got = gpu_batch(glib, w, tmp_path / "g.ctl")
    print("pencils", glib.jur_stat(b"pencils"), "points", glib.jur_stat(b"scattering_points"), flush=True)
    check(got, oracle_batch(olib, w, tmp_path / "o.ctl"))


@pytest.mark.parametrize("fan", [16, 10])
def test_single_scattering_fan(glib, olib, tmp_path, fan):
    """Fewer secondary pencils per scattering point (option fan; 16 is the
    aerosol configuration's count) on the same rays as the 28-pencil test,
    against the oracle with the same fan."""
    w = synth.make("aerosol_limb_8192", tmp_path, nray=64)
    w = w.select(rays=_crossing(w, 20.0, 4, 0), nd=8)
    glib.jur_set_option(b"fan", fan)
    olib.oracle_set_fan(fan)
    try:
        got = gpu_batch(glib, w, tmp_path / "g.ctl")
        pencils = glib.jur_stat(b"pencils")
        points = glib.jur_stat(b"scattering_points")
        ref = oracle_batch(olib, w, tmp_path / "o.ctl")
    finally:
        glib.jur_set_option(b"fan", 28)
        olib.oracle_set_fan(28)
    print("fan", fan, "pencils", pencils, "points", points, flush=True)
    # primaries + a fan and one solar pencil per scattering point
    assert pencils <= w.nray + (fan + 1) * points
    check(got, ref)


def test_unsupported_fan_exits(tmp_path):
    import subprocess
    import sys
    code = (
        "import sys; sys.path[:0] = %r\n"
        "from pathlib import Path\n"
        "from pyjurassic import lib as L, synth; from helpers import gpu_batch\n"
        "lib = L.load(); lib.jur_set_option(b'fan', 34)\n"
        "w = synth.make('aerosol_limb_8192', Path(%r), nray=2)\n"
        "gpu_batch(lib, w, Path(%r) / 'g.ctl')\n" % (sys.path[:4], str(tmp_path), str(tmp_path)))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "Unsupported number of secondary pencils" in r.stdout
