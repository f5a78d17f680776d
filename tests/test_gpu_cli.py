"""The formod command (jurassic-scatter_amd/bin/formod, the reference CLI's
arguments and keys) on the reference examples: output files against the
goldens, as the reference's own regression scripts compare them, and
TASK c (per-emitter contributions and extinction only) against the oracle."""
import ctypes as C
import subprocess

import numpy as np
import pytest

import oracle_ctypes
from conftest import EXAMPLES, ROOT
from helpers import EXAMPLE_CASES, obs_rows
from pyjurassic import jtypes as J

pytestmark = pytest.mark.gpu

FORMOD = ROOT / "jurassic-scatter_amd" / "bin" / "formod"


def run_cli(args, tmp_path):
    r = subprocess.run([str(FORMOD)] + [str(a) for a in args], cwd=EXAMPLES, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout[-500:] + r.stderr[-500:]
    return r


@pytest.mark.parametrize("case", sorted(EXAMPLE_CASES))
def test_cli_vs_golden(case, tmp_path):
    ctlf, obsf, atmf, aerof, golden = EXAMPLE_CASES[case]
    out = tmp_path / "rad.tab"
    args = [ctlf, obsf, atmf, out] + (["AEROFILE", aerof] if aerof else [])
    run_cli(args, tmp_path)
    g = np.array(obs_rows(EXAMPLES / "org" / golden))
    m = np.array(obs_rows(out))
    assert m.shape == g.shape
    np.testing.assert_allclose(m[:, :10], g[:, :10], rtol=1e-5, atol=1e-6)  # geometry, tangent points
    np.testing.assert_allclose(m[:, 10], g[:, 10], rtol=1e-4)             # radiance
    np.testing.assert_allclose(m[:, 11], g[:, 11], rtol=1e-4, atol=1e-5)  # transmittance


def test_cli_task_c(tmp_path):
    """TASK c writes rad.tab plus rad.tab.<emitter> (only that emitter, no
    extinction, no N2/O2 continua) and rad.tab.EXTINCT (no emitters)."""
    out = tmp_path / "rad.tab"
    run_cli(["clear-air.ctl", "org/obs.tab", "atm_pwin_0_80.tab", out, "TASK", "c"], tmp_path)
    olib = oracle_ctypes.load()
    import os
    cwd = os.getcwd()
    os.chdir(EXAMPLES)
    try:
        ctl, obs, atm = J.ctl_t(), J.obs_t(), J.atm_t()
        argc, argv = J.argv("clear-air.ctl")
        olib.oracle_read_ctl(argc, argv, C.byref(ctl))
        olib.oracle_read_obs(None, b"org/obs.tab", C.byref(ctl), C.byref(obs))
        olib.oracle_read_atm(None, b"atm_pwin_0_80.tab", C.byref(ctl), C.byref(atm))
        ctl.ctm_n2 = 0
        ctl.ctm_o2 = 0
        aero = J.aero_t()
        names = [ctl.emitter[ig].value.decode() for ig in range(ctl.ng)]
        for ig, name in enumerate(names + ["EXTINCT"]):
            a2 = J.atm_t.from_buffer_copy(atm)
            for iw in range(ctl.nw):
                for ip in range(a2.np):
                    a2.k[iw][ip] = 0.0 if name != "EXTINCT" else a2.k[iw][ip]
            for ig2 in range(ctl.ng):
                if ig2 != ig:
                    for ip in range(a2.np):
                        a2.q[ig2][ip] = 0.0
            o2 = J.obs_t.from_buffer_copy(obs)
            olib.oracle_formod(C.byref(ctl), C.byref(a2), C.byref(o2), C.byref(aero))
            ref = np.array([o2.rad[0][i] for i in range(o2.nr)])
            got = np.array(obs_rows(tmp_path / ("rad.tab." + name)))[:, 10]
            # EXTINCT of this example is zero up to rounding (no extinction, no
            # emitters): 1e-9 is 1e-7 of the example's radiances
            np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-9, err_msg=name)
    finally:
        os.chdir(cwd)
