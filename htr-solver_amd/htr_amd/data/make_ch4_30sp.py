"""Writes ch4_30sp_species.json: the species data (molar masses, 2-range NASA-7
polynomials, transport parameters) of the 30-species methane mixture of
src/Mixtures/CH4_30SpMix.hpp, read from the reference header as text.  Only
the species block is taken (the reference's reactions are not: the batched-
chemistry workload uses a seeded synthetic mechanism, htr_amd.mechanism).

    python make_ch4_30sp.py /root/reference/src/Mixtures/CH4_30SpMix.hpp
"""
import json
import os
import re
import sys

ATOM = 1e-10          # src/Utils/constants.h: Angstrom to meter
DTOCM = 3.33564e-30   # src/Utils/constants.h: Debye to Coulomb meter
GEOM = {"SpeciesGeom_Atom": 0, "SpeciesGeom_Linear": 1, "SpeciesGeom_NonLinear": 3}


def num(expr):
    expr = expr.strip().rstrip(",").replace("ATom", repr(ATOM)).replace("DToCm", repr(DTOCM))
    if not re.fullmatch(r"[0-9eE+\-*/. ()]+", expr):
        raise ValueError("unexpected expression %r" % expr)
    return float(eval(expr, {"__builtins__": {}}))


def field(block, label):
    m = re.search(r"/\*\s*" + re.escape(label) + r"\s*=\s*\*/\s*([^\\\n]*)", block)
    if not m:
        raise KeyError(label)
    return m.group(1).strip()


def main(path):
    text = open(path).read()
    species = []
    for m in re.finditer(r"#define i(\w+) (\d+)\n#define\s+(\w+) \{(.*?)\n\s*\}\n", text, re.S):
        name, idx, block = m.group(1), int(m.group(2)), m.group(4)
        arr = lambda lab: [num(x) for x in field(block, lab).strip(" {},").split(",") if x.strip()]
        species.append({
            "name": name, "index": idx,
            "W": num(field(block, "W")),
            "TSwitch1": num(field(block, "TSwitch1")), "TSwitch2": num(field(block, "TSwitch2")),
            "TMin": num(field(block, "TMin")), "TMax": num(field(block, "TMax")),
            "cpM": arr("cpM"), "cpL": arr("cpL"),
            "geom": GEOM[field(block, "Geom").rstrip(", ")],
            "sigma": num(field(block, "sigma")), "kbOveps": num(field(block, "kbOveps")),
            "mu": num(field(block, "mu")), "alpha": num(field(block, "alpha")),
            "Z298": num(field(block, "Z298")),
        })
    species.sort(key=lambda s: s["index"])
    assert [s["index"] for s in species] == list(range(30)), [s["index"] for s in species]
    for s in species:
        assert len(s["cpM"]) == 9 and len(s["cpL"]) == 9, s["name"]
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ch4_30sp_species.json")
    with open(out, "w") as fh:
        json.dump({"source": "src/Mixtures/CH4_30SpMix.hpp (species block)", "nPoly": 2, "species": species},
                  fh, indent=1)
    print("wrote", out, len(species), "species")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/Mixtures/CH4_30SpMix.hpp")
