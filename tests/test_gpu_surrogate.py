"""The learned scattering source in the forward model: training pairs from
the plain secondary-pencil method, a network trained on them on the GPU,
and the forward model with the network in place of the plain angular
integral at the primaries' scattering points, compared with the plain
method on rays the network has not seen; the fallback path (every point
outside the training hull) reproduces the plain method exactly."""
import ctypes as C

import numpy as np
import pytest

from helpers import gpu_batch
from pyjurassic import lib as L
from pyjurassic import synth

pytestmark = pytest.mark.gpu

F = C.POINTER(C.c_float)


@pytest.fixture(scope="module")
def glib():
    return L.load()


@pytest.fixture(scope="module")
def case(tmp_path_factory):
    """Aerosol layer at 15-20 km, single scattering: 384 training rays and
    128 held-out rays crossing the layer, 8 channels."""
    tmp = tmp_path_factory.mktemp("sur")
    w = synth.make("aerosol_limb_8192", tmp, nray=2048)
    cross = np.where((w.geom[4] > 0) & (w.geom[4] < 20.0))[0]
    train = w.select(rays=cross[:384], nd=8)
    test = w.select(rays=cross[384:512], nd=8)
    return tmp, train, test


@pytest.fixture(scope="module")
def case2(tmp_path_factory):
    """Cirrus at 9-12 km, scattering depth 2: 240 limb + 240 nadir training
    rays and 12 + 12 held-out rays through the cloud, 8 channels."""
    tmp = tmp_path_factory.mktemp("sur2")
    w = synth.make("cirrus_4096", tmp, nray=1024)
    limb = np.where((w.geom[4] > 0) & (w.geom[4] < 12.0))[0]
    nadir = np.where((w.geom[4] == 0) & (w.geom[6] != 0))[0]
    train = w.select(rays=np.concatenate([limb[:240], nadir[:240]]), nd=8)
    test = w.select(rays=np.concatenate([limb[240:252], nadir[240:252]]), nd=8)
    return tmp, train, test


def run(glib, w, path, surrogate, margin=0.25):
    glib.jur_set_option(b"surrogate", surrogate)
    glib.jur_set_option(b"hull_margin", margin)
    out = gpu_batch(glib, w, path)
    return out, glib.jur_stat(b"scatter_ms")


def train_on(glib, w, path, nd, steps, batch, lr=1e-3):
    glib.jur_set_option(b"collect", 1)
    run(glib, w, path, 0)
    glib.jur_set_option(b"collect", 0)
    n = glib.jur_sca_collect(None, None, 0)
    npt = int(glib.jur_stat(b"scattering_points0"))
    assert n == npt * nd
    feat = np.zeros((n, 24), dtype=np.float32)
    jth = np.zeros(n, dtype=np.float32)
    assert glib.jur_sca_collect(feat.ctypes.data_as(F), jth.ctypes.data_as(F), n) == n
    assert np.all(np.isfinite(feat)) and np.all(jth > 0)
    y = (np.log(jth) - feat[:, 12]).astype(np.float32)  # ln(J / B(T))
    glib.jur_mlp_init(7)
    glib.jur_mlp_set_data(feat.ctypes.data_as(F), y.ctypes.data_as(F), n, 1, 1)
    hist = np.zeros(steps, dtype=np.float32)
    glib.jur_mlp_train(steps, batch, C.c_float(lr), hist.ctypes.data_as(F))
    # a quarter as many steps at a tenth of the rate
    tail = np.zeros(steps // 4, dtype=np.float32)
    glib.jur_mlp_train(steps // 4, batch, C.c_float(lr / 10), tail.ctypes.data_as(F))
    print("pairs %d, loss %.4f -> %.5f -> %.5f" % (n, hist[:20].mean(), hist[-20:].mean(), tail[-20:].mean()),
          flush=True)
    return n


def compare(glib, tmp, test):
    (rp, tp, _, _), ms_plain = run(glib, test, tmp / "p.ctl", 0)
    pencils_plain = glib.jur_stat(b"pencils")
    (rs, ts, _, _), ms_sur = run(glib, test, tmp / "s.ctl", 1)
    nsur = glib.jur_stat(b"surrogate_points")
    nfb = glib.jur_stat(b"fallback_points")
    rel = np.abs(rs / rp - 1)
    print("surrogate points %d, fallback %d; max rel rad %.2e, mean %.2e; plain %.1f ms (%d pencils), "
          "surrogate %.1f ms (%d pencils): %.1fx" % (nsur, nfb, rel.max(), rel.mean(), ms_plain, pencils_plain,
                                                    ms_sur, glib.jur_stat(b"pencils"), ms_plain / ms_sur),
          flush=True)
    glib.jur_set_option(b"surrogate", 0)
    np.testing.assert_allclose(ts, tp, rtol=1e-6, atol=1e-7)  # transmittance has no source term
    return rel, nsur, nfb, ms_plain / ms_sur


def test_surrogate_single_scattering(glib, case):
    """Depth 1: radiance within 0.5 % of the plain method on held-out rays
    (the call time there is dominated by the primaries and solar pencils)."""
    tmp, train, test = case
    n = train_on(glib, train, tmp / "t.ctl", 8, 3000, 4096)
    assert n > 10000
    rel, nsur, nfb, speed = compare(glib, tmp, test)
    assert nsur > 0.9 * (nsur + nfb)
    assert rel.max() <= 5e-3


def test_surrogate_depth2_speed(glib, case2):
    """Depth 2: the network replaces two levels of secondary fans; the
    scattering call is >= 10x faster than the plain method.  Trained on
    ~90k pairs, the network's radiances differ from the plain method's by
    <= 0.05 % on average and <= 0.5 % at most on held-out rays (the largest
    differences: nadir views of the thick cloud in window channels, where
    the radiance follows J almost one to one)."""
    tmp, train, test = case2
    train_on(glib, train, tmp / "t.ctl", 8, 4000, 2048)
    rel, nsur, nfb, speed = compare(glib, tmp, test)
    assert nsur > 0.5 * (nsur + nfb)
    assert rel.mean() <= 5e-4
    assert rel.max() <= 5e-3
    assert speed >= 10.0


def test_fallback_is_plain(glib, case):
    """With an empty hull every point falls back: the result is the plain
    method's, bit for bit."""
    tmp, _, test = case
    (rp, tp, _, _), _ = run(glib, test, tmp / "p2.ctl", 0)
    glib.jur_mlp_init(8)
    glib.jur_mlp_set_data(np.zeros((32, 24), np.float32).ctypes.data_as(F), np.zeros(32, np.float32).ctypes.data_as(F),
                          32, 1, 1)
    glib.jur_mlp_train(1, 32, C.c_float(1e-3), None)
    (rs, ts, _, _), _ = run(glib, test, tmp / "s2.ctl", 1, margin=-1e30)
    assert glib.jur_stat(b"surrogate_points") == 0 and glib.jur_stat(b"fallback_points") > 0
    np.testing.assert_array_equal(rs, rp)
    np.testing.assert_array_equal(ts, tp)
    glib.jur_set_option(b"surrogate", 0)
    glib.jur_set_option(b"hull_margin", 0.25)
