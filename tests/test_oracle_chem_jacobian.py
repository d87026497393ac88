"""Pins of the oracle's analytic chemistry Jacobian (CPU only).

The reference forms the Rosenbrock Jacobian of the chemistry right-hand side
(src/prometeo_chem.hpp:66-88) by central differences
(src/Utils/math_utils.hpp:249-270).  The analytic Jacobian
(o_chem_jacobian_analytic) is the exact derivative of the same function; it
has no reference counterpart, so it is pinned by the mathematics:

  exact differences  the central difference of the right-hand side with the
                     temperature solved to machine precision at every
                     perturbed state (h = 1e-6 |x_j|; truncation and
                     rounding leave ~1e-7 of the row maximum) equals the
                     analytic matrix to 1e-5 of the row maximum.  A dropped
                     temperature term, a sign of the backward-rate
                     derivative, a missing third-body or clamped-species
                     term each leave O(1) errors in some row.  The
                     reference's own differences miss this matrix by 1e-4
                     (air) to 2e-3 (30 species): its temperature Newton
                     stops before a small perturbation moves T, so its
                     temperature columns are noise at that level.
  same integral      the Rosenbrock integration with either Jacobian lands
                     within the step tolerance of the converged solution
                     (tolerance 1e-13) to the same accuracy, and the two
                     agree within ten times the 1e-10 step tolerance.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from htr_amd import cases
from hotpath import air_state, complete_state, dt_cfl, oracle_fields

L = O.lib()
dp = O.dptr


def _T_exact(m, e, Y, T0):
    T = T0
    for _ in range(60):
        f = e - L.o_GetInternalEnergy(m.ptr, T, dp(Y))
        h = 1e-6 * T
        cv = (L.o_GetInternalEnergy(m.ptr, T + h, dp(Y)) - L.o_GetInternalEnergy(m.ptr, T - h, dp(Y))) / (2 * h)
        dT = f / cv
        T += dT
        if abs(dT) < 1e-15 * T:
            break
    return T


def _rhs_exact(m, x, T0):
    """The chemistry right-hand side (Conserved_t_old = 0) with T converged to
    machine precision, from the oracle's point functions."""
    ns = m.nSpec
    rho = x[:ns].sum()
    Y = np.clip(x[:ns] / rho, 1e-60, 1.0)
    MixW = L.o_GetMolarWeightFromYi(m.ptr, dp(Y))
    u = x[ns:ns + 3] / rho
    e = x[ns + 3] / rho - 0.5 * (u * u).sum()
    T = _T_exact(m, e, Y, T0)
    P = L.o_GetPFromRhoAndT(m.ptr, rho, MixW, T)
    w = np.zeros(ns)
    L.o_GetProductionRates(m.ptr, dp(w), rho, P, T, dp(Y))
    return np.concatenate([w, np.zeros(4)]), T


def _case(kind):
    if kind == "chem30":
        n = (2, 2, 2)
        st, meta = cases.chem30_box(2, seed=0)
    else:
        n = (6, 6, 6)
        st, meta = air_state(6, seed=0, composition="per-seed")
    st, om = complete_state(st, meta, n, sensors=False)
    return st, meta, n, om


@pytest.mark.parametrize("kind", ["air", "chem30"])
def test_analytic_jacobian_equals_exact_differences(kind):
    st, meta, n, om = _case(kind)
    g, f = oracle_fields(st, om.nSpec, n)
    N = om.nSpec + 4
    cons = f["Conserved"].reshape(N, -1)
    Tc = f["temperature"].ravel()
    worst, worst_ref = 0.0, 0.0
    for c in range(0, cons.shape[1], max(1, cons.shape[1] // 4)):
        x = cons[:, c].copy()
        _, T0 = _rhs_exact(om, x, Tc[c])
        Je = np.zeros((N, N))
        for j in range(N):
            h = 1e-6 * abs(x[j]) + 1e-14
            xp, xm = x.copy(), x.copy()
            xp[j] += h
            xm[j] -= h
            Je[:, j] = (_rhs_exact(om, xp, T0)[0] - _rhs_exact(om, xm, T0)[0]) / (xp[j] - xm[j])
        Ja, _ = O.chem_jacobian(om, x, np.zeros(N), T0, O.JAC_ANALYTIC)
        Jf, _ = O.chem_jacobian(om, x, np.zeros(N), T0, O.JAC_FD)
        sc = np.abs(Je).max(axis=1, keepdims=True)
        sc[sc == 0] = 1.0
        worst = max(worst, float((np.abs(Ja - Je) / sc).max()))
        worst_ref = max(worst_ref, float((np.abs(Jf - Je) / sc).max()))
        assert np.all(Ja[om.nSpec:] == 0.0)
    print("%s: analytic vs exact differences %.2e, reference differences vs exact %.2e" % (kind, worst, worst_ref))
    assert worst <= 1e-5


@pytest.mark.parametrize("kind", ["air", "chem30"])
def test_analytic_jacobian_integration(kind):
    st, meta, n, om = _case(kind)
    dt = dt_cfl(st, meta, n) if kind == "air" else 1e-5
    ns = om.nSpec
    res = {}
    for lab, jac, tol in (("fd", O.JAC_FD, 1e-10), ("analytic", O.JAC_ANALYTIC, 1e-10),
                          ("converged", O.JAC_ANALYTIC, 1e-13)):
        g, f = oracle_fields(st, ns, n)
        f["Conserved_t"] = 0.0
        nst = np.zeros(f["rho"].shape, np.int32)
        O.chem_update(g, f, dt, om, nst, jacobian=jac, tol=tol)
        assert np.all(nst > 0)
        res[lab] = f["Conserved"][:ns].copy()
    e_fd = float(np.abs(res["fd"] - res["converged"]).max())
    e_an = float(np.abs(res["analytic"] - res["converged"]).max())
    d = float(np.abs(res["analytic"] - res["fd"]).max())
    print("%s: |fd - converged| %.2e, |analytic - converged| %.2e, |analytic - fd| %.2e" % (kind, e_fd, e_an, d))
    assert e_an <= 2.0 * e_fd + 1e-10
    assert d <= 1e-9
