"""The packaged network (jurassic-scatter_amd/data/surrogate_cirrus.bin,
trained on 4.18 M pairs of the plain method over the cirrus training set)
on held-out rays of the 16,384-ray cirrus set: 64 limb and 64 nadir rays
through the cloud, all 32 channels, scattering depth 2, by day.  Against the
plain secondary-pencil method on the same rays: radiance within 0.5 % at
most, the scattering call at least 10x faster."""
from pathlib import Path

import numpy as np
import pytest

from helpers import gpu_batch
from pyjurassic import lib as L
from pyjurassic import synth

pytestmark = pytest.mark.gpu

WEIGHTS = Path(__file__).resolve().parents[1] / "jurassic-scatter_amd" / "data" / "surrogate_cirrus.bin"


def test_packaged_network_cirrus_16384(tmp_path):
    glib = L.load()
    assert glib.jur_mlp_load(str(WEIGHTS).encode()) == 0
    w = synth.make("cirrus_16384", tmp_path, nray=2048)
    limb = np.where((w.geom[4] > 0) & (w.geom[4] < 12.0))[0][:64]
    nadir = np.where((w.geom[4] == 0) & (w.geom[6] != 0))[0][:64]
    t = w.select(rays=np.concatenate([limb, nadir]))
    print("cirrus_16384 subset: %d rays (%d limb, %d nadir) x %d channels" % (t.nray, len(limb), len(nadir),
                                                                             len(t.nu)), flush=True)
    res = {}
    try:
        for sur in (0, 1, 0, 1):  # the second call of each is timed
            glib.jur_set_option(b"surrogate", sur)
            out = gpu_batch(glib, t, tmp_path / "c.ctl")
            res[sur] = (out, glib.jur_stat(b"scatter_ms"), glib.jur_stat(b"pencils"),
                        glib.jur_stat(b"surrogate_points"), glib.jur_stat(b"fallback_points"))
    finally:
        glib.jur_set_option(b"surrogate", 0)
    (rp, tp, _, _), ms_p, pen_p, _, _ = res[0]
    (rs, ts, _, _), ms_s, pen_s, nsur, nfb = res[1]
    rel = np.abs(rs / rp - 1)
    print("plain %.1f ms (%d pencils), surrogate %.1f ms (%d pencils): %.1fx; surrogate points %d, fallback %d; "
          "max rel rad %.2e, mean %.2e" % (ms_p, pen_p, ms_s, pen_s, ms_p / ms_s, nsur, nfb, rel.max(), rel.mean()),
          flush=True)
    np.testing.assert_allclose(ts, tp, rtol=1e-6, atol=1e-7)
    assert nsur > 0.5 * (nsur + nfb)
    assert rel.max() <= 5e-3
    assert ms_p / ms_s >= 10.0
