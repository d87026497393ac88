"""The oracle against the golden vectors of the reference's own unit tests.

Fixtures: tests/golden/*.json, transcribed by tests/golden/make_golden.py from
unitTests/{variablesTest,chemTest,mixTest,cflTest,mathUtilsTest,metricTest}.
Tolerances are the ones the reference tests use.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
RGAS = 8.3144598


def load(name):
    with open(os.path.join(GOLD, name)) as fh:
        return json.load(fh)


def test_ref_mix_setup_consistent():
    assert O.lib().o_mix_sizeof() > 0


def airmix_variables_refs(v):
    """Reference scales of variablesTest/chemTest (their XiRef from Y_ref)."""
    MixWRef = 1.0 / (v["YN2Ref"] / 28.0134e-3 + v["YO2Ref"] / (2 * 15.999e-3))
    xref = {"O2": MixWRef * v["YO2Ref"] / (2 * 15.999e-3), "N2": MixWRef * v["YN2Ref"] / 28.0134e-3}
    m = O.Mix.airmix(v["PRef"], v["TRef"], v["LRef"], xref)
    rhoRef = v["PRef"] * MixWRef / (v["R"] * v["TRef"])
    return m, MixWRef, rhoRef


def test_variables_golden():
    v = load("variables.json")
    m, MixWRef, rhoRef = airmix_variables_refs(v)
    PRef, TRef, LRef = v["PRef"], v["TRef"], v["LRef"]
    uRef = math.sqrt(PRef / rhoRef)
    muRef = math.sqrt(PRef * rhoRef) * LRef
    lamRef = math.sqrt(PRef * rhoRef) * LRef * v["R"] / MixWRef
    DiRef = math.sqrt(PRef / rhoRef) * LRef
    g = O.Grid(3, 3, 3)
    f = O.Fields(g, 5)
    f["pressure"] = v["P"] / PRef
    f["temperature"] = v["T"] / TRef
    f["MolarFracs"] = np.array(v["Xi"])[:, None, None, None]
    f["velocity"] = (np.array(v["v"]) / uRef)[:, None, None, None]
    O.update_properties(g, f, m)
    rel = lambda a, e: abs(a / e - 1.0)
    assert np.all(rel(f["rho"] * rhoRef, v["eRho"]) < 1e-8)
    assert np.all(rel(f["mu"] * muRef, v["eMu"]) < 1e-8)
    assert np.all(rel(f["lam"] * lamRef, v["eLam"]) < 1e-8)
    for i in range(5):
        assert np.all(rel(f["Di"][i] * DiRef, v["eDi"][i]) < 1e-8)
    assert np.all(rel(f["SoS"] * uRef, v["eSoS"]) < 1e-8)
    O.update_conserved(g, f, m)
    for i in range(9):
        assert np.all(rel(f["Conserved"][i], v["eConserved"][i]) < 1e-8), i
    # primitive recovery from the conserved state, Newton guess Tres
    f2 = O.Fields(g, 5)
    f2["Conserved"] = f["Conserved"]
    f2["temperature"] = v["Tres"] / TRef
    O.update_primitive(g, f2, m)
    assert np.all(rel(f2["pressure"] * PRef, v["P"]) < 1e-8)
    assert np.all(rel(f2["temperature"] * TRef, v["T"]) < 1e-8)
    for i in range(5):
        assert np.all(rel(f2["MolarFracs"][i], v["Xi"][i]) < 1e-8)
    for i in range(3):
        assert np.all(rel(f2["velocity"][i] * uRef, v["v"][i]) < 1e-8)


def test_chem_golden():
    c = load("chem.json")
    m, MixWRef, rhoRef = airmix_variables_refs(c)
    PRef, TRef, LRef = c["PRef"], c["TRef"], c["LRef"]
    eRef = PRef / rhoRef
    wiRef = math.sqrt(PRef * rhoRef) / LRef
    Y = np.array(c["Yi"])
    MixW = O.molar_weight_from_Y(m, Y)
    X = O.molar_fractions(m, MixW, Y)
    e = O.call("o_GetInternalEnergy", m, c["T"] / TRef, Y)
    rho = O.call("o_GetRho", m, c["P"] / PRef, c["T"] / TRef, MixW)
    g = O.Grid(2, 2, 2)
    f = O.Fields(g, 5)
    f["pressure"] = c["P"] / PRef
    f["temperature"] = c["T"] / TRef
    f["MassFracs"] = Y[:, None, None, None]
    f["MolarFracs"] = X[:, None, None, None]
    f["velocity"] = 0.0
    f["rho"] = rho
    cons = np.concatenate([rho * Y, [0.0, 0.0, 0.0, rho * e]])
    f["Conserved"] = cons[:, None, None, None]
    O.chem_source(g, f, m)
    for i in range(9):
        err = f["Conserved_t"][i] * wiRef - c["Eprod"][i]
        if c["Eprod"][i] != 0:
            err = err / c["Eprod"][i]
        assert np.all(np.abs(err) < c["tol_explicit"]), (i, f["Conserved_t"][i].ravel()[0] * wiRef)
    dt = c["dt_seconds"] * math.sqrt(PRef / rhoRef) / LRef
    O.chem_update(g, f, dt, m)
    for i in range(9):
        if i == 8:
            val = f["Conserved"][i] * rhoRef * eRef
        else:
            val = f["Conserved"][i] * rhoRef
        err = val - c["Cres"][i]
        if c["Cres"][i] != 0:
            err = err / c["Cres"][i]
        assert np.all(np.abs(err) < c["tol_implicit"]), (i, val.ravel()[0])
    # the explicit rates are copied to Conserved_t_old by UpdateChemistry
    assert np.array_equal(f["Conserved_t_old"], f["Conserved_t"])


def _mix_checks(m, E, refs, T, P, tol, constprop):
    R = 8.3144598
    ok = lambda a, e: abs(a / e - 1.0) < tol
    okd = lambda a, e: abs((a - e) / (e + 1e-60)) < tol
    Y = np.array(E["EYi"])
    MixW = O.molar_weight_from_Y(m, Y)
    assert ok(MixW, E["EMixW"])
    X = O.molar_fractions(m, MixW, Y)
    for i in range(len(Y)):
        assert ok(X[i], E["EXi"][i])
    assert ok(O.molar_weight_from_X(m, X), MixW)
    Y2 = O.mass_fractions(m, MixW, X)
    assert np.all(np.abs(Y2 / Y - 1) < tol)
    rho = O.call("o_GetRho", m, P / refs["PRef"], T / refs["TRef"], MixW) * refs["rhoRef"]
    assert ok(rho, E["Erho"])
    cp = O.call("o_GetHeatCapacity", m, T / refs["TRef"], Y2) * refs["CpRef"]
    assert ok(cp, E["Ecp"])
    h = O.call("o_GetEnthalpy", m, T / refs["TRef"], Y2) * refs["eRef"]
    assert ok(h, E["Eh"])
    for i in range(len(Y)):
        hs = O.call("o_GetSpeciesEnthalpy", m, i, T / refs["TRef"]) * refs["eRef"]
        assert ok(hs, E["Ehsp"][i])
    e = O.call("o_GetInternalEnergy", m, T / refs["TRef"], Y2) * refs["eRef"]
    assert ok(e, E["Eh"] - P / E["Erho"])
    T1 = O.call("o_GetTFromInternalEnergy", m, e / refs["eRef"], (T + 100.0) / refs["TRef"], Y2) * refs["TRef"]
    assert ok(T1, T)
    assert O.call("o_isValidInternalEnergy", m, e / refs["eRef"], Y2) == 1
    assert O.call("o_isValidInternalEnergy", m, -1.0e60, Y2) == 0
    assert ok(O.call("o_GetTFromRhoAndP", m, rho / refs["rhoRef"], MixW, P / refs["PRef"]) * refs["TRef"], T)
    assert ok(O.call("o_GetPFromRhoAndT", m, rho / refs["rhoRef"], MixW, T / refs["TRef"]) * refs["PRef"], P)
    assert ok(O.call("o_GetViscosity", m, T / refs["TRef"], X) * refs["muRef"], E["Emu"])
    assert ok(O.call("o_GetHeatConductivity", m, T / refs["TRef"], X) * refs["lamRef"], E["Elam"])
    gamma = O.call("o_GetGamma", m, T / refs["TRef"], MixW, Y2)
    assert ok(gamma, E["Egamma"])
    assert ok(O.call("o_GetSpeedOfSound", m, T / refs["TRef"], gamma, MixW) * refs["uRef"], E["Esos"])
    dif = O.diffusivity(m, P / refs["PRef"], T / refs["TRef"], MixW, X)
    for i in range(len(Y)):
        assert okd(dif[i] * refs["DiRef"], E["Edif"][i])
    prod = O.production_rates(m, rho / refs["rhoRef"], P / refs["PRef"], T / refs["TRef"], Y2)
    for i in range(len(Y)):
        assert okd(prod[i] * refs["wiRef"], E["Eprod"][i]), (i, prod[i] * refs["wiRef"])
    assert ok(O.call("o_Getdpde", m, rho / refs["rhoRef"], gamma) * refs["dpdeRef"], E["Erho"] * (E["Egamma"] - 1.0))
    dp = O.dpdrhoi(m, gamma, T / refs["TRef"], Y2)
    for i in range(len(Y)):
        assert okd(dp[i] * refs["dpdrhoiRef"], E["Edpdrho"][i])


def test_mix_golden_airmix():
    g = load("mix.json")
    E = g["AirMix"]
    cfg = g["AirMix_json"]
    xref = {s["Name"]: s["MolarFrac"] for s in cfg["XiRef"]["Species"]}
    m = O.Mix.airmix(cfg["PRef"], cfg["TRef"], cfg["LRef"], xref)
    R = 8.3144598
    PRef, TRef, LRef, MixWRef = E["PRef"], E["TRef"], E["LRef"], E["MixWRef"]
    rhoRef = PRef * MixWRef / (R * TRef)
    refs = dict(PRef=PRef, TRef=TRef, rhoRef=rhoRef, eRef=PRef / rhoRef, uRef=math.sqrt(PRef / rhoRef),
                CpRef=R / MixWRef, muRef=math.sqrt(PRef * rhoRef) * LRef,
                lamRef=math.sqrt(PRef * rhoRef) * LRef * R / MixWRef, DiRef=math.sqrt(PRef / rhoRef) * LRef,
                wiRef=math.sqrt(PRef * rhoRef) / LRef, dpdeRef=rhoRef, dpdrhoiRef=PRef / rhoRef)
    E = dict(E)
    E["Erho"] = g["P"] * E["EMixW"] / (R * g["T"])
    _mix_checks(m, E, refs, g["T"], g["P"], g["tol"], False)


def test_mix_golden_constprop():
    g = load("mix.json")
    E = dict(g["ConstPropMix"])
    cfg = g["ConstPropMix_json"]
    m = O.Mix.constprop_mix(cfg["gasConstant"], cfg["gamma"], cfg["prandtl"], 0,
                            cfg["viscosityModel"]["Visc"])
    R = cfg["gasConstant"]
    T, P = g["T"], g["P"]
    E.update(EYi=[1.0], EXi=[1.0], EMixW=8.3144598 / R, Erho=P / (R * T), Ecp=1.005025e+03)
    E.update(Eh=E["Ecp"] * T, Ehsp=[E["Ecp"] * T], Emu=5.0e-3, Elam=E["Ecp"] * 5.0e-3 / 0.71,
             Egamma=1.4, Esos=math.sqrt(1.4 * R * T), Edif=[0.0], Eprod=[0.0], Edpdrho=[R * T])
    one = dict(PRef=1.0, TRef=1.0, rhoRef=1.0, eRef=1.0, uRef=1.0, CpRef=1.0, muRef=1.0, lamRef=1.0,
               DiRef=1.0, wiRef=1.0, dpdeRef=1.0, dpdrhoiRef=1.0)
    _mix_checks(m, E, one, T, P, g["tol"], True)


def test_cfl_golden():
    c = load("cfl.json")
    m = O.Mix.constprop_mix(c["gasConstant"], c["gamma"], 0.71, 0, 0.0)
    g = O.Grid(2, 2, 2)
    base = O.Fields(g, 1)
    base["m_d0"] = 1.0; base["m_d1"] = 1.0; base["m_d2"] = 1.0
    base["temperature"] = 0.0; base["MassFracs"] = 1.0; base["MolarFracs"] = 1.0
    base["rho"] = c["rho"]
    cases = []
    f = base.copy(); f["velocity"] = c["velocity"]; f["SoS"] = c["SoS"]; cases.append(f)
    f = base.copy(); f["mu"] = c["mu"]; cases.append(f)
    f = base.copy(); f["lam"] = c["lam"]; cases.append(f)
    f = base.copy(); f["Di"] = c["Di"]; cases.append(f)
    for f, e in zip(cases, c["expected"]):
        s = O.max_spectral_radius(g, f, m)
        assert abs(s / e - 1.0) < c["tol"], (s, e)


def test_math_utils_golden():
    t = load("math_utils.json")
    x = O.lu_solve(np.array(t["A"]), np.array(t["b"]))
    assert np.all(np.abs(x - np.array(t["x"])) < t["lu_tol"])
    x, nsteps = O.rosenbrock_testprob(t["rb_x0"], t["rb_dtTry"], t["rb_DelT"])
    assert nsteps > 0
    for xi, e in zip(x, t["rb_expected"]):
        assert abs(1.0 - xi / e) < t["rb_tol"], (x, t["rb_expected"])


NODE = ["Std_node", "L_S_node", "Lp1_S_node", "Lp2_S_node", "Rm3_S_node", "Rm2_S_node", "Rm1_S_node",
        "R_S_node", "L_C_node", "Lp1_C_node", "Rm2_C_node", "Rm1_C_node", "R_C_node"]


def test_operator_tables_golden():
    ops = load("operators.json")
    L = O.lib()
    checked = 0
    for name, d in ops.items():
        t = NODE.index(name)
        for k, v in d["ReconPlus"].items():
            assert L.o_coeff_ReconPlus(t, int(k)) == v, (name, k); checked += 1
        for k, v in d["ReconMinus"].items():
            assert L.o_coeff_ReconMinus(t, int(k)) == v, (name, k); checked += 1
        for k, v in d["CoeffsPlus"].items():
            assert L.o_coeff_CoeffsPlus(t, int(k)) == v, (name, k); checked += 1
        for k, v in d["CoeffsMinus"].items():
            assert L.o_coeff_CoeffsMinus(t, int(k)) == v, (name, k); checked += 1
        for k, v in d["Interp"].items():
            assert L.o_coeff_Interp(t, int(k)) == v, (name, k); checked += 1
        for k, v in d["Grad"].items():
            assert L.o_coeff_Grad(t, int(k)) == v, (name, k); checked += 1
        for k, v in d["KennedyCoeff"].items():
            assert L.o_coeff_KennedyCoeff(t, int(k)) == v, (name, k); checked += 1
        if "KennedyOrder" in d:
            assert L.o_coeff_KennedyOrder(t) == d["KennedyOrder"], name; checked += 1
        for slot, off in d["Cp"].items():
            k = ["cm2", "cm1", "cp1", "cp2", "cp3"].index(slot)
            assert L.o_coeff_Cp(t, k) == off, (name, slot); checked += 1
    assert checked > 600


LIBHTR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "htr-solver_amd",
                      "libhtr_amd.so")


@pytest.mark.skipif(not os.path.exists(LIBHTR), reason="library not built")
def test_library_stencil_tables_golden():
    """The product library's own tables (the host copy it uploads to the
    device) against the reference's operator tables, entry by entry."""
    import ctypes
    import htr_amd
    L = htr_amd.lib()
    ops = load("operators.json")
    v = ctypes.c_double(0.0)

    def get(table, node, k):
        assert L.htr_stencil_table(table, node, k, ctypes.byref(v)) == 0
        return v.value
    checked = 0
    for name, d in ops.items():
        t = NODE.index(name)
        for table, key in ((1, "ReconPlus"), (2, "ReconMinus"), (3, "CoeffsPlus"), (4, "CoeffsMinus"),
                           (5, "Interp"), (6, "Grad"), (8, "KennedyCoeff")):
            for k, val in d[key].items():
                assert get(table, t, int(k)) == val, (name, key, k)
                checked += 1
        if "KennedyOrder" in d:
            assert get(7, t, 0) == d["KennedyOrder"], name
            checked += 1
        for slot, off in d["Cp"].items():
            k = ["cm2", "cm1", "cp1", "cp2", "cp3"].index(slot)
            assert get(0, t, k) == off, (name, slot)
            checked += 1
    for k in range(14):
        assert get(10, 0, k) == O.lib().o_coeff_p10(k)
    assert checked > 600
