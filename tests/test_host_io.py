"""Host-side file I/O of the GPU library (no GPU needed): the control,
atmosphere, observation and aerosol readers fill the reference-layout
structs byte-identically to the oracle's readers, the writers produce the
same text, and read_shape / copy_aero follow the reference semantics."""
import ctypes as C

import numpy as np
import pytest

import oracle_ctypes
from helpers import EXAMPLE_CASES
from pyjurassic import jtypes as J
from pyjurassic import lib as L

NSHAPE = 10000  # jurassic.h


@pytest.fixture(scope="module")
def glib():
    return L.load()


@pytest.fixture(scope="module")
def olib():
    return oracle_ctypes.load()


def read_all(lib, prefix, case):
    ctlf, obsf, atmf, aerof, _ = EXAMPLE_CASES[case]
    ctl, obs, atm, aero = J.ctl_t(), J.obs_t(), J.atm_t(), J.aero_t()
    extra = ["AEROFILE", aerof] if aerof else []
    argc, argv = J.argv(ctlf, *extra)
    getattr(lib, prefix + "read_ctl")(argc, argv, C.byref(ctl))
    getattr(lib, prefix + "read_obs")(None, obsf.encode(), C.byref(ctl), C.byref(obs))
    getattr(lib, prefix + "read_atm")(None, atmf.encode(), C.byref(ctl), C.byref(atm))
    if aerof:
        getattr(lib, prefix + "read_aero")(None, aerof.encode(), C.byref(ctl), C.byref(aero))
    return ctl, obs, atm, aero


@pytest.mark.parametrize("case", sorted(EXAMPLE_CASES))
def test_readers_match_oracle(glib, olib, in_examples, case):
    a = read_all(glib, "", case)
    b = read_all(olib, "oracle_", case)
    for x, y, name in zip(a, b, ["ctl", "obs", "atm", "aero"]):
        assert bytes(x) == bytes(y), name


@pytest.mark.parametrize("case", ["clear", "aero1"])
def test_writers_match_oracle(glib, olib, in_examples, tmp_path, case):
    ctl, obs, atm, aero = read_all(glib, "", case)
    for i in range(obs.nr):
        for d in range(ctl.nd):
            obs.rad[d][i] = 0.0123456789 * (i + 1) + d
            obs.tau[d][i] = 0.5 / (i + 1)
    glib.write_obs(None, str(tmp_path / "g_obs.tab").encode(), C.byref(ctl), C.byref(obs))
    olib.oracle_write_obs(None, str(tmp_path / "o_obs.tab").encode(), C.byref(ctl), C.byref(obs))
    assert (tmp_path / "g_obs.tab").read_text() == (tmp_path / "o_obs.tab").read_text()
    # atmosphere: written and read back (%g keeps 6 significant digits)
    glib.write_atm(None, str(tmp_path / "g_atm.tab").encode(), C.byref(ctl), C.byref(atm))
    atm2 = J.atm_t()
    glib.read_atm(None, str(tmp_path / "g_atm.tab").encode(), C.byref(ctl), C.byref(atm2))
    assert atm2.np == atm.np
    n = atm.np
    for f in ["z", "p", "t"]:
        np.testing.assert_allclose(np.array(getattr(atm2, f)[:n]), np.array(getattr(atm, f)[:n]), rtol=1e-5)
    for g in range(ctl.ng):
        np.testing.assert_allclose(np.array(atm2.q[g][:n]), np.array(atm.q[g][:n]), rtol=1e-5)


def test_read_shape(glib, in_examples):
    x = (C.c_double * NSHAPE)()
    y = (C.c_double * NSHAPE)()
    n = C.c_int()
    glib.read_shape(b"tab/boxcar_792.0000.filt", x, y, C.byref(n))
    ref = np.loadtxt("tab/boxcar_792.0000.filt", comments="#")
    assert n.value == len(ref)
    np.testing.assert_array_equal(np.array(x[:n.value]), ref[:, 0])
    np.testing.assert_array_equal(np.array(y[:n.value]), ref[:, 1])


def test_copy_aero(glib, in_examples):
    ctl, _, _, aero = read_all(glib, "", "aero1")
    aero.nl = 3  # layers are set up by get_opt_prop
    aero.beta_e[0][0] = 1.5
    aero.p[1][0][7] = 2.5
    dst = J.aero_t()
    glib.copy_aero(C.byref(ctl), C.byref(dst), C.byref(aero), 0)
    assert bytes(dst) == bytes(aero)
    glib.copy_aero(C.byref(ctl), C.byref(dst), C.byref(aero), 1)
    assert dst.nm == aero.nm and dst.nl == aero.nl
    for im in range(dst.nm):
        assert (dst.top_mod[im], dst.nn[im], dst.rr[im], dst.ss[im]) == (0, 0, 0, 1)
    assert dst.beta_e[0][0] == 0 and dst.p[1][0][7] == 0 and dst.nmod[0] == 0


def test_tab2bin_round_trip(tmp_path):
    """The tab2bin command (reference src/tab2bin.c): the example's binary
    tables written as ASCII rows ("p T u eps", every digit kept) and
    converted back give the original .bin files byte for byte; run on the
    host (no GPU call)."""
    import subprocess
    from conftest import EXAMPLES, ROOT
    from test_gpu_formod import _bin_table
    for g in ("CO2", "H2O", "O3"):
        lines = ["# p T u eps"]
        for p, t, u, e in _bin_table(EXAMPLES / "tab" / ("boxcar_792.0000_%s.bin" % g)):
            lines += ["%.17g %.17g %.9g %.9g" % (p, t, u[k], e[k]) for k in range(len(u))]
        (tmp_path / ("boxcar_792.0000_%s.tab" % g)).write_text("\n".join(lines) + "\n")
    r = subprocess.run([str(ROOT / "jurassic-scatter_amd" / "bin" / "tab2bin"), str(EXAMPLES / "clear-air.ctl"),
                        "TBLBASE", str(tmp_path / "boxcar")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout[-500:] + r.stderr[-500:]
    for g in ("CO2", "H2O", "O3"):
        name = "boxcar_792.0000_%s.bin" % g
        assert (tmp_path / name).read_bytes() == (EXAMPLES / "tab" / name).read_bytes(), name
