"""Tabulated aerosol optics ("Ext" model type, reference src/scatter.c:330-418,
512-579) in the oracle.

The reference ships no such table, so the restatement is pinned by what the
rule fixes: each channel takes the table row nearest its wavenumber (the
lower row on a tie, the end rows beyond the table), modes of one layer add
up, NN = 0 selects the table, other NN values (but 1) contribute nothing, and
NN = 1 and the "Const" type stop with the reference's "Implement me!"."""
import ctypes as C
import subprocess
import sys

import numpy as np
import pytest

import oracle_ctypes
from pyjurassic import jtypes as J

NTHETA = 181
NUS = (650.0, 750.0, 780.0, 1000.0, 800.0)
ROW_OF = (0, 0, 1, 2, 1)  # nearest of 700 / 800 / 900; 750 is a tie -> lower


@pytest.fixture(scope="module")
def olib():
    return oracle_ctypes.load()


def write_table(path, scale):
    rows = []
    for i, nu in enumerate((700.0, 800.0, 900.0)):
        ph = [scale * (i + 1) * (1 + 0.01 * a) for a in range(NTHETA)]
        rows.append(" ".join("%.10g" % v for v in [nu, scale * nu / 1000, scale * nu / 2000] + ph))
    path.write_text("# a comment line is skipped\n" + "\n".join(rows) + "\n")


def setup(olib, tmp_path, modes):
    """modes: (top, bottom, type, table scale or None, NN) per mode."""
    (tmp_path / "c.ctl").write_text("NG = 1\nEMITTER[0] = CO2\nND = %d\n" % len(NUS) +
                                    "".join("NU[%d] = %g\n" % (i, v) for i, v in enumerate(NUS)) +
                                    "SCA_N = %d\nSCA_MULT = 1\n" % len(modes))
    lines = []
    for k, (top, bot, typ, scale, nn) in enumerate(modes):
        f = tmp_path / ("t%d.tab" % k)
        if scale is not None:
            write_table(f, scale)
        lines.append("%g %g 0.1 %s %s %g 1.0 1.5" % (top, bot, typ, f, nn))
    (tmp_path / "aero.tab").write_text("\n".join(lines) + "\n")
    ctl, aero = J.ctl_t(), J.aero_t()
    argc, argv = J.argv(tmp_path / "c.ctl")
    olib.oracle_read_ctl(argc, argv, C.byref(ctl))
    olib.oracle_read_aero(None, str(tmp_path / "aero.tab").encode(), C.byref(ctl), C.byref(aero))
    return ctl, aero


def test_nearest_row_and_mode_sums(olib, tmp_path):
    ctl, aero = setup(olib, tmp_path, [(20, 15, "Ext", 1.0, 0), (20, 15, "ext", 3.0, 0), (20, 15, "Ext", 2.0, 5),
                                       (10, 8, "EXT", 1.0, 0)])
    olib.oracle_get_opt_prop(C.byref(ctl), C.byref(aero))
    assert aero.nl == 2 and list(aero.nmod[:2]) == [3, 1]
    nus = (700.0, 800.0, 900.0)
    for id, r in enumerate(ROW_OF):
        nu = nus[r]
        # layer 0: scale 1 + scale 3 (the NN = 5 mode adds nothing)
        assert aero.beta_e[0][id] == pytest.approx(4 * nu / 1000, rel=1e-12)
        assert aero.beta_s[0][id] == pytest.approx(4 * nu / 2000, rel=1e-12)
        assert aero.beta_a[0][id] == pytest.approx(4 * nu / 2000, rel=1e-12)
        assert aero.p[0][id][90] == pytest.approx(4 * (r + 1) * 1.9, rel=1e-12)
        assert aero.beta_e[1][id] == pytest.approx(nu / 1000, rel=1e-12)
        assert aero.p[1][id][0] == pytest.approx(r + 1, rel=1e-12)


@pytest.mark.parametrize("typ,nn", [("Ext", 1), ("Const", 0.3)])
def test_unimplemented_variants_exit(olib, tmp_path, typ, nn):
    code = (
        "import sys; sys.path[:0] = %r\n"
        "import ctypes as C, oracle_ctypes; from pathlib import Path\n"
        "from test_oracle_ext_optics import setup\n"
        "lib = oracle_ctypes.load()\n"
        "ctl, aero = setup(lib, Path(%r), [(20, 15, %r, 1.0, %r)])\n"
        "lib.oracle_get_opt_prop(C.byref(ctl), C.byref(aero))\n" % (sys.path[:4], str(tmp_path), typ, nn))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Implement me!" in r.stdout
