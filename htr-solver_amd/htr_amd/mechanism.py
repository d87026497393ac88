"""Mechanism descriptions for htr_amd.Mixture.from_mechanism (and the oracle's
Mix.from_mechanism): flat species arrays in the layout of
htr_mix_create_species plus a list of reactions.

synthetic_30(seed): the 30 species of the reference's CH4_30SpMix
(thermodynamics and transport from src/Mixtures/CH4_30SpMix.hpp, extracted to
data/ch4_30sp_species.json by data/make_ch4_30sp.py) with a seeded synthetic
mechanism of 180 reversible elementary reactions: each draws 2 reactants and
2 products without replacement (4 distinct species, stoichiometric
coefficients 1), A log-uniform in [1e8, 1e14] cm^3 mol^-1 s^-1 (the usual unit
of mechanism files; stored in SI, m^3 mol^-1 s^-1, as the reference's
reactions are, i.e. times 1e-6), the temperature exponent n
uniform in [-1, 1] and Ea uniform in [0, 2e5] J/mol, with n and Ea/R stored as
float as the reference's Arrhenius coefficients are (src/Mixtures/Reaction.hpp).
A species quadruple whose standard Gibbs energy of reaction exceeds 10 RT in
magnitude at any of 1500, 3000, 5000, 7000 K is drawn again: with random
species the reverse rate k_f/K_c otherwise reaches 1e20 and more times the
forward rate, and the cell ODE is no longer integrable at the solver's
tolerance.
"""
import json
import os

import numpy as np

RGAS = 8.3144598
GIBBS_T = (1500.0, 3000.0, 5000.0, 7000.0)   # temperatures of the equilibrium filter
GIBBS_MAX = 10.0                             # bound on |Delta G / RT| there
_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "ch4_30sp_species.json")


def ch4_30_species():
    with open(_DATA) as fh:
        return json.load(fh)["species"]


def species_arrays(species):
    ns = len(species)
    W = np.array([s["W"] for s in species])
    cpT = np.array([[s["TSwitch1"], s["TSwitch2"], s["TMin"], s["TMax"]] for s in species], np.float32).ravel()
    nPoly = np.full(ns, 2, np.int32)
    cpH = np.zeros(9 * ns)
    cpM = np.array([s["cpM"] for s in species]).ravel()
    cpL = np.array([s["cpL"] for s in species]).ravel()
    geom = np.array([s["geom"] for s in species], np.int32)
    dc = np.array([[s["sigma"], s["kbOveps"], s["mu"], s["alpha"], s["Z298"]] for s in species]).ravel()
    return dict(W=W, cpT=cpT, nPoly=nPoly, cpH=cpH, cpM=cpM, cpL=cpL, geom=geom, dc=dc,
                names=[s["name"] for s in species])


def gibbs_rt(s, T):
    """G/(RT) of a species from its 2-range NASA-7 polynomial."""
    c = s["cpM"] if T > s["TSwitch1"] else s["cpL"]
    a, b1, b2 = c[2:7], c[7], c[8]
    h = a[0] + a[1] * T / 2 + a[2] * T ** 2 / 3 + a[3] * T ** 3 / 4 + a[4] * T ** 4 / 5 + b1 / T
    sR = a[0] * np.log(T) + a[1] * T + a[2] * T ** 2 / 2 + a[3] * T ** 3 / 3 + a[4] * T ** 4 / 4 + b2
    return h - sR


def synthetic_30(seed=0, nreac=180, PRef=1.0e5, TRef=500.0, LRef=1.0, xref=None):
    species = ch4_30_species()
    mech = species_arrays(species)
    ns = len(species)
    rng = np.random.default_rng(seed)
    reactions = []
    g = np.array([[gibbs_rt(sp, T) for sp in species] for T in GIBBS_T])
    for _ in range(nreac):
        while True:
            idx = rng.choice(ns, size=4, replace=False)
            dg = g[:, idx[2]] + g[:, idx[3]] - g[:, idx[0]] - g[:, idx[1]]
            if np.all(np.abs(dg) <= GIBBS_MAX):
                break
        A = float(10.0 ** rng.uniform(8.0, 14.0)) * 1.0e-6
        n = float(np.float32(rng.uniform(-1.0, 1.0)))
        EovR = float(np.float32(rng.uniform(0.0, 2.0e5) / RGAS))
        reactions.append(dict(kind=0, A=A, n=n, EovR=EovR, has_backward=True,
                              educts=[(int(idx[0]), 1.0), (int(idx[1]), 1.0)],
                              products=[(int(idx[2]), 1.0), (int(idx[3]), 1.0)]))
    xref = xref or {"N2": 0.79, "O2": 0.21}
    XiRef = np.full(ns, 1.0e-60)
    for k, v in xref.items():
        XiRef[mech["names"].index(k)] = v
    mech.update(reactions=reactions, PRef=PRef, TRef=TRef, LRef=LRef, XiRef=XiRef, seed=seed)
    return mech
