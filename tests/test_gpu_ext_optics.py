"""Tabulated aerosol optics ("Ext", reference src/scatter.c:512-579) in the
GPU library: layer optics of get_opt_prop identical to the oracle's, and the
clear example's limb rays through an Ext layer at 10-15 km, extinction only
and single scattering, against the oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle_ctypes
from pyjurassic import jtypes as J
from pyjurassic import lib as L
from test_oracle_ext_optics import setup, write_table

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def glib():
    return L.load()


@pytest.fixture(scope="module")
def olib():
    return oracle_ctypes.load()


def test_layer_optics_identical(glib, olib, tmp_path):
    modes = [(20, 15, "Ext", 1.0, 0), (20, 15, "Ext", 3.0, 0), (10, 8, "Ext", 2.0, 0)]
    out = []
    for lib, prefix in ((olib, "oracle_"), (glib, "")):
        d = tmp_path / (prefix or "gpu")
        d.mkdir()
        ctl, aero = setup(olib, d, modes)
        # the GPU library reads the same files with its own readers
        if not prefix:
            ctl, aero = J.ctl_t(), J.aero_t()
            argc, argv = J.argv(d / "c.ctl")
            lib.read_ctl(argc, argv, C.byref(ctl))
            lib.read_aero(None, str(d / "aero.tab").encode(), C.byref(ctl), C.byref(aero))
        getattr(lib, prefix + "get_opt_prop")(C.byref(ctl), C.byref(aero))
        out.append(aero)
    o, g = out
    for f in ("beta_e", "beta_s", "beta_a"):
        np.testing.assert_array_equal(np.array(getattr(g, f))[:2, :5], np.array(getattr(o, f))[:2, :5], err_msg=f)
    np.testing.assert_array_equal(np.array(g.p)[:2, :5], np.array(o.p)[:2, :5])


@pytest.mark.parametrize("mult", [0, 1])
def test_formod_through_ext_layer(glib, olib, in_examples, tmp_path, mult):
    table = tmp_path / "ext.tab"
    rows = []
    for i, nu in enumerate((700.0, 790.0, 800.0, 900.0)):
        ph = [0.5 / np.pi * (1 + 0.6 * np.cos(np.pi * a / 180)) for a in range(181)]
        rows.append(" ".join("%.10g" % v for v in [nu, 0.02 + 0.01 * i, 0.015 + 0.005 * i] + ph))
    table.write_text("\n".join(rows) + "\n")
    (tmp_path / "aero.tab").write_text("15 10 0.5 Ext %s 0 1 1.5\n" % table)
    keys = ("SCA_N", 1, "SCA_MULT", mult)
    res = []
    for lib, prefix in ((olib, "oracle_"), (glib, "")):
        ctl, obs, atm, aero = J.ctl_t(), J.obs_t(), J.atm_t(), J.aero_t()
        argc, argv = J.argv("clear-air.ctl", *keys)
        getattr(lib, prefix + "read_ctl")(argc, argv, C.byref(ctl))
        getattr(lib, prefix + "read_obs")(None, b"org/obs.tab", C.byref(ctl), C.byref(obs))
        getattr(lib, prefix + "read_atm")(None, b"atm_pwin_0_80.tab", C.byref(ctl), C.byref(atm))
        getattr(lib, prefix + "read_aero")(None, str(tmp_path / "aero.tab").encode(), C.byref(ctl), C.byref(aero))
        getattr(lib, prefix + "get_opt_prop")(C.byref(ctl), C.byref(aero))
        getattr(lib, prefix + "formod")(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(aero))
        res.append((np.array([obs.rad[0][i] for i in range(obs.nr)]), np.array([obs.tau[0][i] for i in range(obs.nr)])))
    (ro, to), (rg, tg) = res
    rel = np.max(np.abs(rg / ro - 1))
    print("SCA_MULT %d: max rel rad %.2e, max |dtau| %.2e" % (mult, rel, np.max(np.abs(tg - to))), flush=True)
    assert rel < 1e-4
    np.testing.assert_allclose(tg, to, atol=1e-5)
