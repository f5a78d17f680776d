"""Write the example atmospheres the reference's regression scripts build.

The reference's shell regressions (examples/clear.sh, extinction1.sh,
aerosol0.sh, aerosol1.sh) first run its `climatology` tool, which picks one
zone of the Remedios climatology (src/climatology.c:152-278: latitude 20 /
month 1 for the default "midl" zone, latitude 65 / month 1 for "pwin") and
interpolates it to integer altitudes (src/climatology.c:282-371) before
writing it with write_atm (src/atmosphere.c:545-598, "%g" columns).  Only the
pwin 0-80 km atmosphere is kept in examples/org/atm.tab; the other runs'
atmospheres are regenerated here from the climatology data file so that the
fixtures examples/org/rad_ext1.tab, rad_aero0.tab and rad_aero1.tab have
their inputs.  At integer altitudes the interpolation returns the
climatology's node values, so the output is the zone's table printed with %g.
The pwin output is checked byte-for-byte against examples/org/atm.tab by
tests/test_oracle_golden.py (that pins this restatement).

Usage: python make_example_atm.py /root/reference/examples/clim_remedios.tab
"""
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
EMITTERS = ["CO2", "H2O", "O3"]


def zone_profiles(climfile, lat, month, zangle=0):
    prof = {}
    for line in Path(climfile).read_text(errors="replace").splitlines():
        tok = line.split()
        if len(tok) < 10 or tok[0].startswith("#"):
            continue
        try:
            lat1, _lat2, month1, _m2, zang1, _z2 = (int(v) for v in tok[1:7])
            alt, mean = float(tok[7]), float(tok[8])
        except ValueError:
            continue
        if lat == lat1 and month == month1 and zangle == zang1:
            prof.setdefault(tok[0], {})[int(alt)] = mean
    return prof


def g(x):
    return "%g" % x


def write_atm(path, prof, zmin, zmax):
    out = ["# $1 = time (seconds since 2000-01-01T00:00Z)\n",
           "# $2 = altitude [km]\n", "# $3 = longitude [deg]\n",
           "# $4 = latitude [deg]\n", "# $5 = pressure [hPa]\n",
           "# $6 = temperature [K]\n"]
    n = 6
    for e in EMITTERS:
        n += 1
        out.append("# $%d = %s volume mixing ratio\n" % (n, e))
    out.append("# $%d = window 0: extinction [1/km]\n" % (n + 1))
    for i, z in enumerate(range(zmin, zmax + 1)):
        if i == 0:
            out.append("\n")
        row = ["%.2f" % 0.0, g(z), "0", "0", g(prof["PRESSURE"][z]),
               g(prof["TEMPERATURE"][z])]
        row += [g(prof[e][z]) for e in EMITTERS] + ["0"]
        out.append(" ".join(row) + "\n")
    Path(path).write_text("".join(out))


def main(climfile, outdir=HERE / "examples"):
    outdir = Path(outdir)
    outdir.mkdir(parents=True, exist_ok=True)
    midl = zone_profiles(climfile, 20, 1)
    pwin = zone_profiles(climfile, 65, 1)
    write_atm(outdir / "atm_midl_0_80.tab", midl, 0, 80)
    write_atm(outdir / "atm_pwin_0_80.tab", pwin, 0, 80)
    write_atm(outdir / "atm_pwin_0_60.tab", pwin, 0, 60)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/examples/clim_remedios.tab")
