"""Regenerate surfparam_tables.inc from the reference's public-domain SurfaceParam.c.

Smoldyn turns surface interaction rates into per-collision probabilities by
interpolating tabulated solutions of the adsorption problem:
lookupirrevadsorb's probability/kappa pairs (source/libSteve/SurfaceParam.c:
504-524) and lookuprevads's grid of reversible-adsorption probabilities
(SurfaceParam.c:566-662).  Parity of those probabilities with the reference
requires the same numbers, so they are extracted verbatim as numbers (no code)
into a data-only include file.  Array names repeat across functions of the
source file, so each table is taken from inside its own function.

Usage: python gen_surfparam_tables.py /root/reference/source/libSteve/SurfaceParam.c
"""
import re
import sys
from pathlib import Path

# (function, array in that function, output name, entries)
TABLES = [("lookupirrevadsorb", "ponlist", "irr_pon", 21),
          ("lookupirrevadsorb", "kappalist", "irr_kappa", 21),
          ("lookuprevads", "xlist", "rev_x", 21),
          ("lookuprevads", "adsprobtable", "rev_adsprob", 21 * 21)]


def function_body(text, fname):
    m = re.search(r"\ndouble\s+%s\s*\(" % fname, text)
    if not m:
        raise SystemExit("function %s not found" % fname)
    end = text.find("\n/* ", m.end())
    return text[m.end():end if end > 0 else len(text)]


def extract(body, name):
    m = re.search(r"double\s+%s\[\]\s*=\s*\{(.*?)\};" % name, body, re.S)
    if not m:
        raise SystemExit("table %s not found" % name)
    body = re.sub(r"//[^\n]*", "", m.group(1))
    vals = [v.strip() for v in body.replace("\n", " ").split(",") if v.strip()]
    for v in vals:
        float(v)
    return vals


def main():
    src = Path(sys.argv[1]).read_text()
    out = ["/* Data only: tabulated adsorption probabilities of Smoldyn's surface",
           "   parameter lookups (public domain, source/libSteve/SurfaceParam.c:504-524,",
           "   566-662).  Generated by gen_surfparam_tables.py; do not edit. */"]
    for fname, arr, name, n in TABLES:
        vals = extract(function_body(src, fname), arr)
        if len(vals) != n:
            raise SystemExit("table %s.%s has %d entries, expected %d" % (fname, arr, len(vals), n))
        out.append("static const double SMOLSP_%s[%d] = {" % (name, n))
        for i in range(0, n, 7):
            out.append("  " + ",".join(vals[i:i + 7]) + ",")
        out.append("};")
    Path(__file__).with_name("surfparam_tables.inc").write_text("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
