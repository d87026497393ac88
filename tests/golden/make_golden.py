"""Writes the golden-vector fixtures of tests/golden/*.json.

The values are the expected numbers that the reference's own unit tests hold
(HTR-solver unitTests/*).  This script only reads those files as text and
transcribes the numbers: no reference code is run or imported.  It needs the
reference checkout at /root/reference; the JSON files it writes are committed
so that the tests never need the checkout.

    python tests/golden/make_golden.py
"""
import json
import os
import re
import sys

REF = os.environ.get("HTR_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))
_NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"


def _read(rel):
    with open(os.path.join(REF, rel)) as fh:
        return fh.read()


def _arith(expr):
    """Evaluates a constant expression made of numbers, + - * / and ( )."""
    expr = expr.strip()
    if not re.fullmatch(r"[0-9eE.+\-*/() ]+", expr):
        raise ValueError("not a constant expression: %r" % expr)
    return float(eval(expr, {"__builtins__": {}}, {}))  # digits and operators only


def _rexpr_values(text):
    """Maps `local NAME = rexpr ... end` to a float or a list of floats."""
    out = {}
    for m in re.finditer(r"local\s+(\w+)\s*=\s*rexpr\s+(.*?)\s+end", text, re.S):
        name, body = m.group(1), m.group(2).strip()
        arr = re.fullmatch(r"array\((.*)\)", body, re.S)
        try:
            if arr:
                items = [s.strip() for s in arr.group(1).split(",")]
                out[name] = [_arith(s) for s in items]
            else:
                out[name] = _arith(body)
        except ValueError:
            continue
    return out


def variables_test():
    v = _rexpr_values(_read("unitTests/variablesTest/variablesTest.rg"))
    keys = ["R", "P", "T", "Xi", "v", "Tres", "eRho", "eMu", "eLam", "eDi", "eSoS",
            "eConserved", "LRef", "PRef", "TRef", "YO2Ref", "YN2Ref"]
    return {k: v[k] for k in keys}


def chem_test():
    t = _read("unitTests/chemTest/chemTest.rg")
    v = _rexpr_values(t)
    keys = ["R", "P", "T", "Yi", "Eprod", "Cres", "LRef", "TRef", "PRef", "YO2Ref", "YN2Ref"]
    out = {k: v[k] for k in keys}
    tol = re.findall(r"fabs\(err\)\s*<\s*(" + _NUM + ")", t)
    out["tol_explicit"], out["tol_implicit"] = float(tol[0]), float(tol[1])
    out["dt_seconds"] = float(re.search(r"var dt = (" + _NUM + r")\*sqrt", t).group(1))
    return out


def mix_test():
    t = _read("unitTests/mixTest/mixTest.rg")
    out = {}
    for name in ("ConstPropMix", "AirMix"):
        start = t.index('(os.getenv("EOS") == "%s")' % name)
        end = t.index("elseif", start)
        block = t[start:end]
        vals = {}
        for m in re.finditer(r"(\w+)\s*=\s*rexpr\s+(.*?)\s+end", block, re.S):
            body = m.group(2).strip()
            arr = re.fullmatch(r"array\((.*)\)", body, re.S)
            try:
                if arr:
                    vals[m.group(1)] = [_arith(s) for s in arr.group(1).split(",")]
                else:
                    vals[m.group(1)] = _arith(body)
            except ValueError:
                continue
        for m in re.finditer(r"local\s+(\w+)\s*=\s*(" + _NUM + r")\s*$", block, re.M):
            vals.setdefault(m.group(1), float(m.group(2)))
        out[name] = vals
    out["T"] = float(re.search(r"local T = (" + _NUM + ")", t).group(1))
    out["P"] = float(re.search(r"local P = (" + _NUM + ")", t).group(1))
    out["tol"] = 1e-3
    for name in ("ConstPropMix", "AirMix"):
        out[name + "_json"] = json.loads(_read("unitTests/mixTest/%s.json" % name))["Flow"]["mixture"]
    return out


def cfl_test():
    t = _read("unitTests/cflTest/cflTest.rg")
    exp = [float(x) for x in re.findall(r"fabs\(s/(" + _NUM + r") - 1\.0\)", t)]
    return {"expected": exp, "gasConstant": 1.0, "gamma": 1.4, "SoS": 10.0,
            "velocity": 1.0, "mu": 10.0, "lam": 10.0, "Di": 10.0, "rho": 1.0, "tol": 1e-3}


def math_utils_test():
    t = _read("unitTests/mathUtilsTest/CPPmathUtilsTest.cc")
    A = [[0.0] * 3 for _ in range(3)]
    for i, j, val in re.findall(r"A\((\d), (\d)\) =\s*(" + _NUM + ");", t):
        A[int(i)][int(j)] = float(val)
    b = [0.0] * 3
    for i, val in re.findall(r"b\[(\d)\] =\s*(" + _NUM + ");", t):
        b[int(i)] = float(val)
    sol = [3.0, 4.0, -5.0]
    rb = [float(x) for x in re.findall(r"x\[\d\]/\(\s*(" + _NUM + r")\)", t)]
    x0 = [float(x) for x in re.findall(r"x\[\d\] = (" + _NUM + ");", t)]
    solve = re.search(r"p->solve\(x, (" + _NUM + "), (" + _NUM + r")\)", t)
    return {"A": A, "b": b, "x": sol, "lu_tol": 1e-12, "rb_x0": x0,
            "rb_dtTry": float(solve.group(1)), "rb_DelT": float(solve.group(2)),
            "rb_expected": rb, "rb_tol": 1e-6}


_OFFS = {"int3d([c])": 0, "[mk_cm2]": -2, "[mk_cm1]": -1, "[mk_cp1]": 1,
         "[mk_cp2]": 2, "[mk_cp3]": 3}


def _check_blocks(text):
    starts = [m.start() for m in re.finditer(r"(?:local )?function check\w+\(r, c,", text)]
    starts.append(len(text))
    for a, b in zip(starts[:-1], starts[1:]):
        yield text[a:b]


def operators_test():
    tables = {}
    for fname in ("operatorsTest_Periodic.rg", "operatorsTest_Staggered.rg",
                  "operatorsTest_Collocated.rg"):
        text = _read("unitTests/metricTest/" + fname)
        for blk in _check_blocks(text):
            nt = re.search(r"r\[c\]\.\[nType\] == (\w+_node)", blk)
            if not nt:
                continue
            node = nt.group(1)
            d = tables.setdefault(node, {"source": [], "ReconPlus": {}, "ReconMinus": {},
                                         "CoeffsPlus": {}, "CoeffsMinus": {}, "Interp": {},
                                         "Grad": {}, "KennedyCoeff": {}, "Cp": {}})
            d["source"].append(fname)
            for k, pat in (("ReconPlus", "GetRecon_Plus"), ("ReconMinus", "GetRecon_Minus")):
                for s, i, val in re.findall(pat + r"\(rexpr r\[c\]\.\[nType\] end, Stencil(\d), (\d)\)\]\s*==\s*([^,]+),", blk):
                    d[k]["%d" % (6 * (int(s) - 1) + int(i))] = _arith(val)
            for k, pat in (("CoeffsPlus", "GetCoeffs_Plus"), ("CoeffsMinus", "GetCoeffs_Minus")):
                for s, val in re.findall(pat + r"\(rexpr r\[c\]\.\[nType\] end, Stencil(\d)\)\]\s*==\s*([^,]+),", blk):
                    d[k]["%d" % (int(s) - 1)] = _arith(val)
            for k, pat in (("Interp", "GetInterp"), ("Grad", "GetGrad"), ("KennedyCoeff", "GetKennedyCoeff")):
                for i, val in re.findall(pat + r"\(rexpr r\[c\]\.\[nType\] end, (\d)\)\]\s*==\s*([^,]+),", blk):
                    d[k][i] = _arith(val)
            ko = re.search(r"GetKennedyOrder\(rexpr r\[c\]\.\[nType\] end\)\]\s*==\s*(\d+)", blk)
            if ko:
                d["KennedyOrder"] = int(ko.group(1))
            for slot, rhs in re.findall(r"\[(cm2|cm1|cp1|cp2|cp3)_d\]\s*==\s*(int3d\(\[c\]\)|\[mk_c[mp]\d\])", blk):
                d["Cp"][slot] = _OFFS[rhs]
    return tables


def main():
    if not os.path.isdir(REF):
        sys.exit("reference checkout not found at %s" % REF)
    fixtures = {
        "variables.json": variables_test(),
        "chem.json": chem_test(),
        "mix.json": mix_test(),
        "cfl.json": cfl_test(),
        "math_utils.json": math_utils_test(),
        "operators.json": operators_test(),
    }
    for name, data in fixtures.items():
        with open(os.path.join(HERE, name), "w") as fh:
            json.dump(data, fh, indent=1, sort_keys=True)
        print("wrote", name)


if __name__ == "__main__":
    main()
