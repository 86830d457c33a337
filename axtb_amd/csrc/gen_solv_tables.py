#!/usr/bin/env python3
"""Generate ``solv_tables.h``: the GFN2-xTB ALPB/GBSA solvent parameters, the D3 van der Waals
radii of the surface model and the Lebedev-Laikov generators of the angular grids.

The tables are data of the published parametrisations (xtb's ``param_alpb_*.fh`` /
``param_gbsa_*.fh`` and the Lebedev-Laikov grids), which axtb keeps as Rust constants. This script
reads those ``.rs`` files as text (it never runs or imports reference code) and writes every
literal as a round-tripping decimal, so the header holds the same bits.

Sources (relative to the reference checkout):
  crates/solvation/src/params.rs        solvent parameter sets, selection tables, aliases, radii
  crates/solvation/src/lebedev_data.rs  grid sizes and Oh generators per grid

Usage: python gen_solv_tables.py [/root/reference] > solv_tables.h   (run by hand; output committed)
"""
import re
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
NELEM = 94


def strip_comments(text):
    return re.sub(r"//[^\n]*", "", text)


def num(s):
    return float(s.strip())


def fmt(x):
    r = repr(float(x))
    return r if ("e" in r or "." in r or "inf" in r or "nan" in r) else r + ".0"


def parse_sets(text):
    """name -> dict of one `pub static NAME: RawParam = RawParam { ... };` per parameter set."""
    sets = {}
    for m in re.finditer(r"pub static ([A-Z0-9_]+): RawParam = RawParam \{(.*?)\n\};", text, re.S):
        body = m.group(2)
        d = {"name": re.search(r'name:\s*"([^"]+)"', body).group(1)}
        for key in ("epsv", "smass", "rhos", "c1", "rprobe", "gshift", "soset", "alpha"):
            d[key] = num(re.search(key + r":\s*([-0-9.eE+]+)", body).group(1))
        for key in ("gamscale", "sx", "tmp"):
            arr = re.search(key + r":\s*\[(.*?)\]", body, re.S).group(1)
            vals = [num(v) for v in arr.split(",") if v.strip()]
            assert len(vals) == NELEM, (m.group(1), key, len(vals))
            d[key] = vals
        sets[m.group(1)] = d
    return sets


def parse_table(text, name):
    body = re.search(r"pub static " + name + r": &\[\(&str, &RawParam\)\] = &\[(.*?)\];", text, re.S).group(1)
    return re.findall(r'\("([^"]+)",\s*&([A-Z0-9_]+)\)', body)


def parse_aliases(text):
    body = re.search(r"pub static NORMALIZE: &\[\(&str, &str\)\] = &\[(.*?)\];", text, re.S).group(1)
    return re.findall(r'\("([^"]+)",\s*"([^"]+)"\)', body)


def parse_array(text, name):
    body = re.search(r"pub static " + name + r": \[f64; NELEM\] = \[(.*?)\];", text, re.S).group(1)
    vals = [num(v) for v in body.split(",") if v.strip()]
    assert len(vals) == NELEM
    return vals


def parse_lebedev(text):
    sizes = [int(v) for v in re.search(r"GRID_SIZE: \[usize; \d+\] = \[(.*?)\];", text, re.S).group(1).split(",")
             if v.strip()]
    body = re.search(r"pub static GRIDS: \[&\[Gen\]; \d+\] = \[(.*)\];", text, re.S).group(1)
    grids = []
    for g in re.finditer(r"&\[(.*?)\]", body, re.S):
        gens = []
        for k, args in re.findall(r"Oh([1-6])\(([^)]*)\)", g.group(1)):
            a = [num(v) for v in args.split(",") if v.strip()]
            k = int(k)
            if k <= 3:
                gens.append((k, 0.0, 0.0, a[0]))
            elif k <= 5:
                gens.append((k, a[0], 0.0, a[1]))
            else:
                gens.append((k, a[0], a[1], a[2]))
        grids.append(gens)
    assert len(grids) == len(sizes)
    npts = {1: 6, 2: 12, 3: 8, 4: 24, 5: 24, 6: 48}
    for s, g in zip(sizes, grids):
        assert sum(npts[k] for k, *_ in g) == s, s
    return sizes, grids


def main():
    ptext = strip_comments((REF / "crates/solvation/src/params.rs").read_text())
    ltext = strip_comments((REF / "crates/solvation/src/lebedev_data.rs").read_text())
    sets = parse_sets(ptext)
    order = list(sets)
    alpb = parse_table(ptext, "GFN2_ALPB")
    gbsa = parse_table(ptext, "GFN2_GBSA")
    aliases = parse_aliases(ptext)
    vdw = parse_array(ptext, "VDW_RAD_D3_AA")
    sizes, grids = parse_lebedev(ltext)
    o = []
    w = o.append
    w("// Generated by gen_solv_tables.py from the reference's solvation tables (data of the published")
    w("// GFN2-xTB ALPB/GBSA parametrisation and the Lebedev-Laikov grids). Do not edit by hand.")
    w("#pragma once")
    w("")
    w("namespace axtb_solv {")
    w("")
    w(f"constexpr int SOLV_NELEM = {NELEM};")
    w("")
    w("// one parameter set: dielectric constant, solvent molar mass (g/mol) and density (g/cm^3), Born")
    w("// radius scaling c1, probe radius (Angstrom), free-energy shift (kcal/mol), Born offset (0.1")
    w("// Angstrom), and per element the surface-tension scaling, the dielectric descreening and the")
    w("// square root of the H-bond strength ((kcal/mol)^1/2)")
    w("struct SolvParamSet {")
    w("  const char* name;")
    w("  double epsv, smass, rhos, c1, rprobe, gshift, soset;")
    w("  double gamscale[SOLV_NELEM], sx[SOLV_NELEM], tmp[SOLV_NELEM];")
    w("};")
    w("")
    w(f"constexpr int SOLV_NSETS = {len(order)};")
    w("static const SolvParamSet SOLV_SETS[SOLV_NSETS] = {")
    for key in order:
        d = sets[key]
        w("    {\"%s\", %s," % (d["name"], ", ".join(fmt(d[k]) for k in
                                               ("epsv", "smass", "rhos", "c1", "rprobe", "gshift", "soset"))))
        for arr in ("gamscale", "sx", "tmp"):
            vals = d[arr]
            w("     {" + ", ".join(fmt(v) for v in vals) + "},")
        w("    },")
    w("};")
    w("")
    w("struct SolvName {")
    w("  const char* name;")
    w("  int set;")
    w("};")
    for tname, tab in (("SOLV_ALPB", alpb), ("SOLV_GBSA", gbsa)):
        w(f"constexpr int {tname}_N = {len(tab)};")
        w(f"static const SolvName {tname}[{tname}_N] = {{")
        for n, s in tab:
            w(f"    {{\"{n}\", {order.index(s)}}},")
        w("};")
    w("")
    w("struct SolvAlias {")
    w("  const char* alias;")
    w("  const char* target;")
    w("};")
    w(f"constexpr int SOLV_NALIAS = {len(aliases)};")
    w("static const SolvAlias SOLV_ALIASES[SOLV_NALIAS] = {")
    for a, t in aliases:
        w(f"    {{\"{a}\", \"{t}\"}},")
    w("};")
    w("")
    w("// D3 van der Waals radii (Angstrom) of the surface model")
    w("static const double SOLV_VDW_RAD_D3_AA[SOLV_NELEM] = {" + ", ".join(fmt(v) for v in vdw) + "};")
    w("")
    w("// Lebedev-Laikov grids: sizes, and per grid its Oh generators {kind 1..6, a, b, weight}")
    w(f"constexpr int LEB_NGRID = {len(sizes)};")
    w("static const int LEB_GRID_SIZE[LEB_NGRID] = {" + ", ".join(str(s) for s in sizes) + "};")
    offs = [0]
    for g in grids:
        offs.append(offs[-1] + len(g))
    w("static const int LEB_GEN_OFF[LEB_NGRID + 1] = {" + ", ".join(str(s) for s in offs) + "};")
    w("struct LebGen {")
    w("  int kind;")
    w("  double a, b, v;")
    w("};")
    w(f"static const LebGen LEB_GEN[{offs[-1]}] = {{")
    for g in grids:
        for k, a, b, v in g:
            w(f"    {{{k}, {fmt(a)}, {fmt(b)}, {fmt(v)}}},")
    w("};")
    w("")
    w("}  // namespace axtb_solv")
    print("\n".join(o))


if __name__ == "__main__":
    main()
