"""The scattering configurations at sizes the oracle needs minutes for,
against its committed fixtures (tests/golden/config, written by
tests/golden/make_config_goldens.py on the same seeded inputs; the inputs'
digest is checked first):

* aerosol_limb_64: 64 limb rays through the aerosol layer at 15-20 km,
  all 32 channels, 5 emitters, single scattering by day, 16 pencils per
  fan (the configuration's count);
* cirrus_8: 4 limb and 4 nadir rays through the cirrus at 9-12 km,
  8 channels, 5 emitters, scattering depth 2 by day, 28 pencils per fan;
* aerosol_limb_64_cfg, cirrus_8_cfg: the same at the configurations' own
  parameters (the aerosol layer's Mie grid of 64 log-spaced radii and 180
  phase angles, option mie_grid; 32 pencils per fan for the cirrus).

Bars as for the other batched tests: radiance within 1e-4 relative,
transmittance within 1e-5 absolute, tangent points within 1e-6 km, LOS
point counts equal.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

from helpers import gpu_batch
from pyjurassic import lib as L

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
import make_config_goldens as G  # noqa: E402

pytestmark = pytest.mark.gpu

FIX = Path(__file__).resolve().parent / "golden" / "config"


@pytest.fixture(scope="module")
def glib():
    return L.load()


@pytest.mark.parametrize("case", list(G.CASES))
def test_config_vs_oracle_fixture(glib, tmp_path, case):
    w, fan, opts = G.case_workload(case, tmp_path)
    ref = np.load(FIX / ("%s.npz" % case))
    assert str(ref["digest"]) == G.digest(w), "the generator no longer makes the fixture's inputs"
    print("%s: %d rays x %d channels x %d emitters, fan %d %s" % (case, w.nray, len(w.nu), len(w.emitters), fan,
                                                                  dict(opts)), flush=True)
    glib.jur_set_option(b"fan", fan)
    for k, v in opts:
        glib.jur_set_option(k.encode(), v)
    try:
        rad, tau, tp, npts = gpu_batch(glib, w, tmp_path / "g.ctl")
    finally:
        glib.jur_set_option(b"fan", 28)
        for k, _ in opts:
            glib.jur_set_option(k.encode(), 0)
    rel = np.abs(rad / ref["rad"] - 1)
    bt = np.abs(bt_of(rad, w.nu) - bt_of(ref["rad"], w.nu))
    print("%s: max rel rad %.3e (mean %.2e), max dBT %.2e K, max abs tau %.2e" %
          (case, rel.max(), rel.mean(), bt.max(), np.abs(tau - ref["tau"]).max()), flush=True)
    np.testing.assert_allclose(rad, ref["rad"], rtol=1e-4)
    np.testing.assert_allclose(tau, ref["tau"], atol=1e-5)
    np.testing.assert_allclose(tp, ref["tp"], atol=1e-6)
    np.testing.assert_array_equal(npts, ref["npts"])


def bt_of(rad, nu):
    c1, c2 = 1.19104259e-8, 1.43877506
    nu = np.asarray(nu)[None, :]
    return c2 * nu / np.log1p(c1 * nu ** 3 / rad)
