"""Copies the reference's CHEMKIN input for the H2/CO, H2 and CH4 skeletal mechanisms
into fixtures.

Data only: the kinetics files of run/kinetic-mechanisms/POLIMI_H2CO_1412 and
POLIMI_H2_1412, and the NASA-7 entries of the species they use, cut out of the
shared POLIMI_TOT_NOX_1412.CKT thermodynamics file (its THERMO header line and
global temperature line, the 4-line entries of those species, END); the kinetics
and thermodynamics files of POLIMI_CH4_SKELETAL_1412 (the mechanism of the
MethaneCoflowFlame/02-skeletal-mechanism tutorial) as they are; the one-step global
mechanisms GLOBAL_CH4_1step and GLOBAL_H2_1step with their thermodynamics cut to
their species.

    python tests/golden/make_chemkin_fixtures.py [/root/reference]
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from laminarsmoke_amd.chemkin import read_kinetics  # noqa: E402


def main(ref="/root/reference"):
    base = os.path.join(ref, "run", "kinetic-mechanisms")
    species = []
    for mech in ("POLIMI_H2CO_1412", "POLIMI_H2_1412"):
        src = os.path.join(base, mech, mech + ".CKI")
        text = open(src).read()
        with open(os.path.join(HERE, mech + ".CKI"), "w") as f:
            f.write(text)
        species += [s for s in read_kinetics(text)[1] if s not in species]
    lines = open(os.path.join(base, "POLIMI_H2CO_1412", "POLIMI_TOT_NOX_1412.CKT")).read().splitlines()
    out = lines[:2]
    i = 2
    while i < len(lines):
        ln = lines[i]
        if len(ln) >= 80 and ln[79] == "1" and ln[:18].split()[0].upper() in species:
            out += lines[i:i + 4]
            i += 4
        else:
            i += 1
    out.append("END")
    with open(os.path.join(HERE, "POLIMI_H2CO_1412_thermo.CKT"), "w") as f:
        f.write("\n".join(out) + "\n")
    print(f"{len(species)} species: {species}")
    # one-step global mechanisms (FORD orders), their thermodynamics cut to their species
    for mech in ("GLOBAL_CH4_1step", "GLOBAL_H2_1step"):
        text = open(os.path.join(base, mech, mech + ".CKI")).read()
        with open(os.path.join(HERE, mech + ".CKI"), "w") as f:
            f.write(text)
        sp = read_kinetics(text)[1]
        lines = open(os.path.join(base, mech, mech + ".CKT")).read().splitlines()
        h = next(i for i, ln in enumerate(lines) if ln.strip().upper().startswith("THERMO"))
        out = lines[h:h + 2]
        i = h + 2
        while i < len(lines):
            ln = lines[i]
            if len(ln) >= 80 and ln[79] == "1" and ln[:18].split()[0].upper() in sp:
                out += lines[i:i + 4]
                i += 4
            else:
                i += 1
        out.append("END")
        with open(os.path.join(HERE, mech + "_thermo.CKT"), "w") as f:
            f.write("\n".join(out) + "\n")
    ch4 = os.path.join(base, "POLIMI_CH4_SKELETAL_1412")
    for src, dst in (("POLIMI_CH4_SKELETAL_1412.CKI", "POLIMI_CH4_SKELETAL_1412.CKI"),
                     ("POLIMI_TOT_NOX_1412.CKT", "POLIMI_CH4_SKELETAL_1412_thermo.CKT")):
        with open(os.path.join(ch4, src)) as f, open(os.path.join(HERE, dst), "w") as g:
            g.write(f.read())


if __name__ == "__main__":
    main(*sys.argv[1:])
