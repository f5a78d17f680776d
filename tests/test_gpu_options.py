"""Forward-model options and edge cases that no reference fixture covers,
GPU library against the oracle on the same inputs: REFRAC = 0, hydrostatic
equilibrium (HYDZ), field-of-view convolution (FOV), brightness temperature
output (WRITE_BBT), the N2 / O2 continua, non-finite input radiances (the
reference's mask), an empty batch, and the error exit for an observer below
the surface."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_ctypes
from conftest import ROOT
from helpers import gpu_batch, oracle_batch
from pyjurassic import jtypes as J
from pyjurassic import lib as L
from pyjurassic import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def glib():
    return L.load()


@pytest.fixture(scope="module")
def olib():
    return oracle_ctypes.load()


def run_example(lib, prefix, extra, nan_ray=None):
    ctl, obs, atm, aero = J.ctl_t(), J.obs_t(), J.atm_t(), J.aero_t()
    argc, argv = J.argv("clear-air.ctl", *extra)
    getattr(lib, prefix + "read_ctl")(argc, argv, C.byref(ctl))
    getattr(lib, prefix + "read_obs")(None, b"org/obs.tab", C.byref(ctl), C.byref(obs))
    getattr(lib, prefix + "read_atm")(None, b"atm_pwin_0_80.tab", C.byref(ctl), C.byref(atm))
    if nan_ray is not None:
        obs.rad[0][nan_ray] = float("nan")
    getattr(lib, prefix + "formod")(C.byref(ctl), C.byref(atm), C.byref(obs), C.byref(aero))
    nr = obs.nr
    rad = np.array([obs.rad[0][i] for i in range(nr)])
    tau = np.array([obs.tau[0][i] for i in range(nr)])
    p = np.array([atm.p[i] for i in range(atm.np)])
    return rad, tau, p


@pytest.mark.parametrize("extra", [["REFRAC", "0"], ["HYDZ", "10"], ["WRITE_BBT", "1"],
                                   ["HYDZ", "0", "WRITE_BBT", "1"]])
def test_option_vs_oracle(glib, olib, in_examples, extra):
    g = run_example(glib, "", extra)
    o = run_example(olib, "oracle_", extra)
    print(extra, "max rel rad %.2e" % np.max(np.abs(g[0] / o[0] - 1)), flush=True)
    if "WRITE_BBT" in extra:
        np.testing.assert_allclose(g[0], o[0], atol=0.01)  # brightness temperature [K]
        assert np.all((g[0] > 150) & (g[0] < 320))
    else:
        np.testing.assert_allclose(g[0], o[0], rtol=1e-4)
    np.testing.assert_allclose(g[1], o[1], atol=1e-5)
    np.testing.assert_allclose(g[2], o[2], rtol=1e-12)  # hydrostatic pressure profile (FP64)


def test_fov_vs_oracle(glib, olib, in_examples, tmp_path):
    fov = tmp_path / "fov.tab"
    dz = np.linspace(-2.0, 2.0, 21)
    fov.write_text("".join("%g %g\n" % (z, np.exp(-0.5 * (z / 0.8) ** 2)) for z in dz))
    g = run_example(glib, "", ["FOV", str(fov)])
    o = run_example(olib, "oracle_", ["FOV", str(fov)])
    np.testing.assert_allclose(g[0], o[0], rtol=1e-4)
    np.testing.assert_allclose(g[1], o[1], atol=1e-5)
    plain = run_example(olib, "oracle_", [])
    assert not np.allclose(o[0], plain[0], rtol=1e-6)  # the convolution did something


def test_nan_mask(glib, olib, in_examples):
    g = run_example(glib, "", [], nan_ray=3)
    o = run_example(olib, "oracle_", [], nan_ray=3)
    assert np.isnan(g[0][3]) and np.isnan(o[0][3])
    keep = np.arange(len(g[0])) != 3
    np.testing.assert_allclose(g[0][keep], o[0][keep], rtol=1e-4)


def test_n2_o2_continua(glib, olib, tmp_path):
    """Channels inside the O2 (1360-1805 cm^-1) and N2 (2120-2605 cm^-1)
    continuum bands."""
    w = synth.make("clear_limb_17", tmp_path, nu=[1400.0, 1600.0, 2300.0, 2500.0])
    got = gpu_batch(glib, w, tmp_path / "g.ctl")
    ref = oracle_batch(olib, w, tmp_path / "o.ctl")
    np.testing.assert_allclose(got[0], ref[0], rtol=1e-4)
    np.testing.assert_allclose(got[1], ref[1], atol=1e-5)
    w2 = w.select()
    w2.extra_ctl = dict(w.extra_ctl, CTM_N2="0", CTM_O2="0")
    off = oracle_batch(olib, w2, tmp_path / "o2.ctl")
    assert np.all(off[1] >= ref[1]) and np.any(off[1] > ref[1] * (1 + 1e-3))  # continua absorb


def test_empty_batch(glib, tmp_path):
    """No rays: the call returns without work (one atmosphere, zero rays)."""
    w = synth.make("clear_batch_65536", tmp_path, nray=4).select(rays=[0])
    w.nray = 0
    w.geom = np.ascontiguousarray(w.geom[:, :0])
    out = gpu_batch(glib, w, tmp_path / "g.ctl")
    assert out[0].shape == (0, 64)


def test_observer_below_surface_exits(tmp_path):
    """The reference stops with "Observer below surface!"; so does the
    library (status flag from the ray-trace kernel, exit code 1)."""
    code = r'''
import sys
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, sys.argv[2])
import numpy as np
from pathlib import Path
from helpers import gpu_batch
from pyjurassic import lib as L, synth
w = synth.make("clear_limb_17", Path(sys.argv[3]))
w.geom[1][:] = -5.0   # observers 5 km below the surface
gpu_batch(L.load(), w, Path(sys.argv[3]) / "g.ctl")
print("no error")
'''
    r = subprocess.run([sys.executable, "-c", code, str(ROOT / "jurassic-scatter_amd"), str(ROOT / "tests"),
                        str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-300:], r.stderr[-300:])
    assert r.returncode == 1
    assert "Observer below surface!" in r.stdout and "no error" not in r.stdout
