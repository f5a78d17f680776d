"""Turn continua_coeffs.dat into a C include file (run at build time).

Writes ``static const double JUR_CTM_<NAME>[n] = {...};`` for every table of
the data file, so the oracle and the GPU library compile in the same numbers.

Usage:  python continua_inc.py <output.inc>
"""
import sys
from pathlib import Path


def tables(path):
    lines = Path(path).read_text().split("\n")
    i = 0
    while i < len(lines):
        head = lines[i].split()
        i += 1
        if not head:
            continue
        name, n = head[0], int(head[1])
        vals = []
        while len(vals) < n:
            vals.extend(lines[i].split())
            i += 1
        yield name, vals


def main(out):
    src = Path(__file__).with_name("continua_coeffs.dat")
    parts = ["/* generated from continua_coeffs.dat -- do not edit */\n"]
    for name, vals in tables(src):
        parts.append("static const double JUR_CTM_%s[%d] = {\n" % (name.upper(), len(vals)))
        for k in range(0, len(vals), 6):
            parts.append("  " + ", ".join(vals[k:k + 6]) + ",\n")
        parts.append("};\n")
    Path(out).write_text("".join(parts))


if __name__ == "__main__":
    main(sys.argv[1])
