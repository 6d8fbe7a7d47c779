"""Regenerate rxnparam_tables.inc from the reference's public-domain rxnparam.c.

The tabulated reduced reaction rates that Smoldyn's rate<->radius conversion
interpolates (yi, yr, yb for numrxnrate; kirr, krev, kb for numrxnrateprob,
source/libSteve/rxnparam.c:58-91 and :162-391) are data: parity of binding
radii with the reference requires the same numbers, so they are extracted
verbatim as numbers (no code) into a data-only include file.

Usage: python gen_rxnparam_tables.py /root/reference/source/libSteve/rxnparam.c
"""
import re
import sys
from pathlib import Path

NAMES = ["yi", "yr", "yb", "kirr", "krev", "kb"]
EXPECTED = {"yi": 31, "yr": 16 * 31, "yb": 11 * 31,
            "kirr": 8 * 31, "krev": 16 * 8 * 31, "kb": 11 * 8 * 31}


def extract(text, name):
    m = re.search(r"double\s+%s\[\]\s*=\s*\{(.*?)\};" % name, text, re.S)
    if not m:
        raise SystemExit("table %s not found" % name)
    body = re.sub(r"//[^\n]*", "", m.group(1))
    vals = [v.strip() for v in body.replace("\n", " ").split(",") if v.strip()]
    for v in vals:
        float(v)
    return vals


def main():
    src = Path(sys.argv[1]).read_text()
    out = ["/* Data only: tabulated reduced rates of the Andrews-Bray reaction-parameter",
           "   solver (public domain, source/libSteve/rxnparam.c:58-91,162-391).",
           "   Generated by gen_rxnparam_tables.py; do not edit. */"]
    for n in NAMES:
        vals = extract(src, n)
        if len(vals) < EXPECTED[n]:
            raise SystemExit("table %s has %d entries, expected >= %d" % (n, len(vals), EXPECTED[n]))
        out.append("static const double SMOLTBL_%s[%d] = {" % (n, len(vals)))
        for i in range(0, len(vals), 8):
            out.append("  " + ",".join(vals[i:i + 8]) + ",")
        out.append("};")
    Path(__file__).with_name("rxnparam_tables.inc").write_text("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
