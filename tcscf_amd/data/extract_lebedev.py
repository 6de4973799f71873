"""Extract Lebedev angular quadrature tables into a compact C include file.

The reference ships the Lebedev-Laikov tables as ``long double`` literals in
``src/integration/data/LebedevData_<order>.hpp`` (ordering {polar, azimuthal,
weight}, weights summing to 4*pi) and converts each entry to the working type
with a plain cast (``src/integration/Lebedev.cpp:23-41``).  This script parses
those literals as x86 80-bit extended values (numpy.longdouble) and rounds them
to binary64 exactly as that cast does, then writes the binary64 values as C
hex-float literals, so every consumer sees bit-identical points and weights.

Usage:  python tcscf_amd/data/extract_lebedev.py [/root/reference]
Writes: tcscf_amd/data/lebedev_tables.inc
"""
import os
import re
import sys

import numpy as np

ORDERS = [3, 5, 7, 9, 11, 13, 15, 17, 19, 21, 23, 25, 27, 29, 31, 35, 41]


def parse(path):
    text = open(path).read()
    m = re.search(r"\[\s*(\d+)\s*\]\s*\[\s*3\s*\]", text)
    npts = int(m.group(1))
    body = text[text.index("{", m.end()) + 1:]
    lits = re.findall(r"([-+]?\d+\.\d+(?:[eE][-+]?\d+)?)L", body)
    assert len(lits) == 3 * npts, (path, len(lits), npts)
    vals = np.array([np.longdouble(s) for s in lits], dtype=np.longdouble)
    return npts, vals.astype(np.float64)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    out = os.path.join(here, "lebedev_tables.inc")
    lines = [
        "/* Lebedev-Laikov angular grids: rows of {polar, azimuthal, weight},",
        " * weights normalised to 4*pi.  Generated by extract_lebedev.py from the",
        " * reference tables (src/integration/data/LebedevData_<order>.hpp); values",
        " * are the binary64 roundings of the 80-bit literals.  Data, not code. */",
    ]
    table, offset = [], 0
    for order in ORDERS:
        npts, vals = parse(os.path.join(ref, "src/integration/data", f"LebedevData_{order}.hpp"))
        table.append((order, npts, offset))
        lines.append(f"static const double kLebedev_{order}[{3 * npts}] = {{")
        for i in range(npts):
            a, b, c = (float(v).hex() for v in vals[3 * i:3 * i + 3])
            lines.append(f"  {a}, {b}, {c},")
        lines.append("};")
        offset += npts
    lines.append("struct LebedevEntry { int order; int npts; const double* data; };")
    lines.append("static const LebedevEntry kLebedevTable[] = {")
    for order, npts, _ in table:
        lines.append(f"  {{ {order}, {npts}, kLebedev_{order} }},")
    lines.append("};")
    lines.append(f"static const int kLebedevTableSize = {len(table)};")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {out}: {len(table)} orders, {offset} points")


if __name__ == "__main__":
    main()
