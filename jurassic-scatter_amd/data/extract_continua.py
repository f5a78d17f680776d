"""Extract the continuum absorption coefficient tables into a plain data file.

The CO2, H2O, N2 and O2 continuum models of the forward model are driven by
tabulated physical coefficients (CKD-type continuum data: CO2 at 296/260/230 K
on a 2 cm^-1 grid, H2O self/foreign on a 10 cm^-1 grid with the 820-960 cm^-1
correction factors, and the N2 / O2 collision-induced band coefficients).
Those numbers are data, not code; the reference ships them inside
src/continua.c (ctmco2 :11-837, ctmh2o :883-1884, ctmn2 :1939-1977,
ctmo2 :2001-2034).  This script reads that file once and writes the numbers,
in the same order, to ``continua_coeffs.dat`` next to it.  Both the oracle and
the GPU library compile that data file in at build time.

Usage:  python extract_continua.py /path/to/reference/src/continua.c
"""
import re
import sys
from pathlib import Path

# (name in the data file, array name in the source, expected length, which
# occurrence of that array name -- ba/betaa/nua appear once for N2, once for O2)
TABLES = [
    ("co2_296", "co2296", 2001, 0),
    ("co2_260", "co2260", 2001, 0),
    ("co2_230", "co2230", 2001, 0),
    ("h2o_self296", "h2o296", 2001, 0),
    ("h2o_self260", "h2o260", 2001, 0),
    ("h2o_foreign", "h2ofrn", 2001, 0),
    ("h2o_xfcrev", "xfcrev", 15, 0),
    ("n2_b", "ba", 98, 0),
    ("n2_beta", "betaa", 98, 0),
    ("n2_nu", "nua", 98, 0),
    ("o2_b", "ba", 90, 1),
    ("o2_beta", "betaa", 90, 1),
    ("o2_nu", "nua", 90, 1),
]


def main(src):
    text = Path(src).read_text(errors="replace")
    out = []
    for name, cname, n, occ in TABLES:
        pat = re.compile(r"static\s+double\s+%s\s*\[\s*(\d+)\s*\]\s*=\s*\{(.*?)\}\s*;" % cname, re.S)
        found = pat.findall(text)
        size, body = found[occ]
        vals = [v for v in re.split(r"[\s,]+", body.strip()) if v]
        assert int(size) == n and len(vals) == n, (name, size, len(vals))
        floats = [float(v) for v in vals]
        out.append("%s %d\n" % (name, n))
        for i in range(0, n, 8):
            out.append(" ".join(repr(x) for x in floats[i:i + 8]) + "\n")
    dst = Path(__file__).with_name("continua_coeffs.dat")
    dst.write_text("".join(out))
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/continua.c")
