#!/usr/bin/env python3
"""Generate ``gfn2_tables.h``: the GFN2-xTB, STO-nG, EEQ-2019 and D4 parameter tables.

The tables are data of the published parametrisations; axtb keeps them as generated Rust
constants. This script reads those ``.rs`` files as text (it never runs or imports reference
code), evaluates each literal in IEEE double precision exactly as a constant initialiser does
(``AATOAU * 0.32``, ``1.0 + 0.1972612``, ``EVTOAU * -10.7``, ...) and writes every value as a
round-tripping decimal, so the header holds the same bits as the reference's compiled constants.

Sources (relative to the reference checkout):
  crates/axtb/src/param/gfn2_data.rs   GFN2-xTB element parameters
  crates/axtb/src/param/sto_data.rs    STO-nG expansions (Stewart)
  crates/axtb/src/param/eeq_data.rs    EEQ-2019 (multicharge)
  crates/axtb/src/param/element_data.rs, crates/base/src/elements.rs   radii, electronegativities
  crates/d4/src/data.rs                D4 reference polarisabilities and charges
  crates/base/src/units.rs             the CODATA-2018 bohr of mctc-lib (MCTC_AATOAU)

Usage: python gen_tables.py [/root/reference] > gfn2_tables.h   (run by hand; output committed)
"""
import math
import re
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")

# mctc-lib's bohr (crates/base/src/units.rs, module `reference`), in the same operation order
PLANCK_2018 = 6.62607015e-34
ELECTRON_MASS_2018 = 9.1093837015e-31
SPEED_OF_LIGHT = 299792458.0
FINE_STRUCTURE_2018 = 7.2973525693e-3
HBAR_2018 = PLANCK_2018 / (2.0 * math.pi)
BOHR_METER_2018 = HBAR_2018 / (ELECTRON_MASS_2018 * SPEED_OF_LIGHT * FINE_STRUCTURE_2018)
MCTC_AATOAU = 1.0 / (BOHR_METER_2018 * 1e10)
EVTOAU = 1.0 / 27.21138505


def strip_comments(text):
    return re.sub(r"//[^\n]*", "", text)


def tables(path):
    """name -> nested list of every `pub const|static NAME: [..] = [ ... ];` array in `path`."""
    text = strip_comments((REF / path).read_text())
    out = {}
    for m in re.finditer(r"pub (?:const|static) ([A-Z0-9_]+)\s*:\s*(\[[^=]*\])\s*=\s*\[", text):
        name = m.group(1)
        start = m.end() - 1
        depth = 0
        for i in range(start, len(text)):
            if text[i] == "[":
                depth += 1
            elif text[i] == "]":
                depth -= 1
                if depth == 0:
                    body = text[start:i + 1]
                    break
        env = {"AATOAU": MCTC_AATOAU, "EVTOAU": EVTOAU, "__builtins__": {}}
        out[name] = eval(body, env)  # literals and the two named factors only
    return out


def flat(v):
    if isinstance(v, list):
        for x in v:
            yield from flat(v=x)
    else:
        yield v


def dims(v):
    d = []
    while isinstance(v, list):
        d.append(len(v))
        v = v[0]
    return d


def emit(name, v, ctype="double"):
    d = dims(v)
    vals = list(flat(v))
    fmt = (lambda x: repr(float(x))) if ctype == "double" else (lambda x: str(int(x)))
    lines = []
    for i in range(0, len(vals), 6):
        lines.append("  " + ", ".join(fmt(x) for x in vals[i:i + 6]) + ",")
    shape = "".join(f"[{n}]" for n in d)
    return f"static const {ctype} {name}{shape} = {{\n" + "\n".join(lines) + "\n};\n"


def main():
    g = tables("crates/axtb/src/param/gfn2_data.rs")
    s = tables("crates/axtb/src/param/sto_data.rs")
    q = tables("crates/axtb/src/param/eeq_data.rs")
    e = tables("crates/base/src/elements.rs")
    ed = tables("crates/axtb/src/param/element_data.rs")
    d4 = tables("crates/d4/src/data.rs")
    out = [
        "// Generated by gen_tables.py from axtb's parameter tables (GFN2-xTB, STO-nG, EEQ-2019,\n"
        "// mctc-lib element data, DFT-D4 reference data; LGPL-3.0-or-later data). Do not edit.\n"
        "// Every value is the IEEE double of the reference's constant expression.\n",
        "#pragma once\n",
        f"#define AXTB_MCTC_AATOAU {MCTC_AATOAU!r}\n",
        f"#define AXTB_EVTOAU {EVTOAU!r}\n",
        "#define AXTB_GFN2_MAX_ELEM 86\n#define AXTB_D4_MAX_REF 7\n#define AXTB_D4_NFREQ 23\n#define AXTB_D4_NSEC 17\n",
    ]
    for name in ["REP_ALPHA", "REP_ZEFF", "HUBBARD_PARAMETER", "SHELL_HUBBARD", "P_HUBBARD_DERIVS", "P_SHPOLY",
                 "P_KCN", "REFERENCE_OCC", "SLATER_EXPONENT", "P_SELFENERGY", "P_DKERNEL", "P_QKERNEL",
                 "P_VCN", "P_RAD"]:
        out.append(emit("GFN2_" + name, g[name]))
    for name in ["NSHELL", "ANG_SHELL", "PRINCIPAL_QUANTUM_NUMBER", "NUMBER_OF_PRIMITIVES"]:
        out.append(emit("GFN2_" + name, g[name], "int"))
    for name in ["P_ALPHA1", "P_ALPHA2", "P_COEFF2", "P_ALPHA3", "P_COEFF3", "P_ALPHA4", "P_COEFF4",
                 "P_ALPHA5", "P_COEFF5", "P_ALPHA6", "P_COEFF6", "P_ALPHA_6S", "P_COEFF_6S", "P_ALPHA_6P",
                 "P_COEFF_6P"]:
        out.append(emit("STO_" + name, s[name]))
    for name in ["EEQ_CHI", "EEQ_ETA", "EEQ_KCNCHI", "EEQ_RAD"]:
        out.append(emit(name, q[name]))
    out.append(emit("EL_PAULING_EN", e["PAULING_EN"]))
    out.append(emit("EL_COVALENT_RAD_2009_AA", e["COVALENT_RAD_2009_AA"]))
    out.append(emit("EL_ATOMIC_RAD_AA", ed["ATOMIC_RAD_AA"]))
    for name in ["CHEMICAL_HARDNESS", "EFFECTIVE_NUCLEAR_CHARGE", "R4_OVER_R2", "REFQ_GFN2", "REFH_GFN2",
                 "REFQ_EEQ", "REFH_EEQ", "HCOUNT", "ASCALE", "REFCOVCN", "REFCN", "ALPHAIW", "SECQ", "SSCALE",
                 "SECCN", "SECAIW"]:
        out.append(emit("D4_" + name, d4[name]))
    for name in ["REFN", "REFSYS"]:
        out.append(emit("D4_" + name, d4[name], "int"))
    sys.stdout.write("\n".join(out))


if __name__ == "__main__":
    main()
