"""Shared test helpers: example cases and text comparison of obs files."""
import ctypes as C
from pathlib import Path

from pyjurassic import jtypes as J

# (control file, obs input, atmosphere, aerosol file or None, golden output)
EXAMPLE_CASES = {
    "clear": ("clear-air.ctl", "org/obs.tab", "atm_pwin_0_80.tab", None, "rad_clear.tab"),
    "ext1": ("extinction1.ctl", "org/obs.tab", "atm_midl_0_80.tab", "aero0.tab", "rad_ext1.tab"),
    "aero0": ("aerosol0.ctl", "org/obs_nadir.tab", "atm_pwin_0_60.tab", "aero0.tab", "rad_aero0.tab"),
    "aero1": ("aerosol1.ctl", "org/obs.tab", "atm_midl_0_80.tab", "aero1.tab", "rad_aero1.tab"),
}


def read_example(lib, prefix, case):
    """Read ctl/obs/atm/aero of an example with `lib`'s readers (prefix
    'oracle_' for the oracle, '' for the GPU library)."""
    ctlf, obsf, atmf, aerof, _ = EXAMPLE_CASES[case]
    ctl, obs, atm, aero = J.ctl_t(), J.obs_t(), J.atm_t(), J.aero_t()
    extra = ["AEROFILE", aerof] if aerof else []
    argc, argv = J.argv(ctlf, *extra)
    getattr(lib, prefix + "read_ctl")(argc, argv, C.byref(ctl))
    getattr(lib, prefix + "read_obs")(None, obsf.encode(), C.byref(ctl), C.byref(obs))
    getattr(lib, prefix + "read_atm")(None, atmf.encode(), C.byref(ctl), C.byref(atm))
    if aerof:
        getattr(lib, prefix + "read_aero")(None, aerof.encode(), C.byref(ctl), C.byref(aero))
        getattr(lib, prefix + "get_opt_prop")(C.byref(ctl), C.byref(aero))
    return ctl, obs, atm, aero


def obs_rows(path):
    """Numeric rows of an obs file (comment and blank lines skipped)."""
    rows = []
    for line in Path(path).read_text().splitlines():
        if line.strip() and not line.startswith("#"):
            rows.append([float(v) for v in line.split()])
    return rows
