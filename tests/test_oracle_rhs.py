"""Closed-form pins of the oracle's flux-divergence and stencil restatement
(CPU only).  Each check is chosen so that a plausible slip in the oracle -- a
wrong pair offset in the Kennedy-Gruber sums, a swapped tangential index, a
sign in the split fluxes, a dropped species term -- fails it:

  uniform state            every scheme and the diffusion flux give exactly 0
  telescoping              sum_p Conserved_t / m_e = 0 on a periodic box
  KG split form            the pure KG divergence equals an independent numpy
                           evaluation of the split form from the pair sums
                           S(a, b) = (rho_a + rho_b)(u_a + u_b)(phi_a + phi_b)
                           (Pirozzoli 2010, the form the reference implements)
  KG consistency           for rho, p const and u const the mass equation
                           reduces to -(rho u) times the 6th-order central
                           derivative of phi, which converges to -rho u phi'
  axis permutation         permuting x and y of the state permutes the result
  diffusion                species fluxes sum to zero in every cell; a linear
                           temperature with constant properties gives a zero
                           interior divergence of the heat flux
  TENO all-kept            with every stencil kept the reconstruction is the
                           6th-order central interpolation [1,-8,37,37,-8,1]/60
  TENO polynomials         data from a quadratic are reproduced exactly
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from htr_amd import cases
from hotpath import air_state, complete_state, oracle_fields, tgv_state

SCHEMES = [O.SCHEME_HYBRID, O.SCHEME_TENO, O.SCHEME_TENOA, O.SCHEME_TENOLAD, O.SCHEME_KG]


def air_uniform(n):
    st, meta = air_state(n, seed=3)
    for k in ("pressure", "temperature"):
        st[k] = np.full_like(st[k], float(np.mean(st[k])))
    st["MolarFracs"] = np.broadcast_to(st["MolarFracs"][:, :1, :1, :1], st["MolarFracs"].shape).copy()
    st["velocity"] = np.zeros_like(st["velocity"])
    st["velocity"][0] = 0.3
    st["velocity"][1] = -0.2
    return st, meta


def rhs(state, meta, n, scheme=None, diffusion_dir=None, dirs=(0, 1, 2), RK=(1.0, 0.0), dt=1e-4):
    st, om = complete_state(state, meta, n)
    g, f = oracle_fields(st, om.nSpec, n)
    f["Conserved_t"] = 0.0
    for d in dirs:
        if scheme is not None:
            O.euler_flux(g, f, d, scheme, RK[0], RK[1], dt, om)
        if diffusion_dir is not None:
            O.diffusion_flux(g, f, d, om)
    return f["Conserved_t"].copy(), st, om


@pytest.mark.parametrize("scheme", SCHEMES)
def test_uniform_state_gives_zero_euler_rhs(scheme):
    n = (8, 8, 8)
    st, meta = air_uniform(8)
    ct, _, _ = rhs(st, meta, n, scheme=scheme)
    assert np.all(ct == 0.0)


def test_uniform_state_gives_zero_diffusion():
    n = (8, 8, 8)
    st, meta = air_uniform(8)
    ct, _, _ = rhs(st, meta, n, diffusion_dir=True)
    assert np.max(np.abs(ct)) == 0.0


@pytest.mark.parametrize("scheme", SCHEMES)
def test_periodic_telescoping(scheme):
    n = (12, 12, 12)
    st, meta = air_state(12, seed=1)
    for d in range(3):
        ct, s2, _ = rhs(st, meta, n, scheme=scheme, dirs=(d,), RK=(0.5, 0.5), dt=1e-3)
        m = s2["m_e%d" % d]
        total = np.sum(ct / m[None], axis=(1, 2, 3))
        scale = np.max(np.abs(ct / m[None]), axis=(1, 2, 3)) * ct[0].size
        assert np.all(np.abs(total) <= 1e-12 * scale), (d, total, scale)


def kg_split_form(st, om, d):
    """Independent numpy evaluation of the KG flux divergence along d on a
    periodic box of standard nodes: F(p+1/2) = sum_l c_l sum_{m<=l} S(p-m, p-m+l+1)."""
    ns = om.nSpec
    ax = 3 - d                                    # array axis of direction d in (comp, z, y, x)
    rho = st["rho"][None]
    u = st["velocity"][d][None]
    Y = st["MassFracs"]
    vel = st["velocity"]
    p = st["pressure"][None]
    Q = st["Conserved"]
    H = (Q[ns + 3][None] + p) / rho
    phi = np.concatenate([Y, vel, H], axis=0)    # (nEq, ...)
    c = [3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0]

    def sh(a, k):                                 # value at p + k
        return np.roll(a, -k, axis=ax)

    F = np.zeros_like(Q)
    P = np.zeros_like(p)
    for l in range(3):
        acc = 0.0
        accp = 0.0
        for m in range(l + 1):
            a, b = -m, -m + l + 1
            acc = acc + (sh(rho, a) + sh(rho, b)) * (-(sh(u, a) + sh(u, b))) * (sh(phi, a) + sh(phi, b))
            accp = accp + sh(p, a) + sh(p, b)
        F += c[l] * acc
        P += c[l] * accp
    F = 0.25 * F
    F[ns + d] -= P[0]
    m_e = st["m_e%d" % d][None]
    return (F - np.roll(F, 1, axis=ax)) * m_e


@pytest.mark.parametrize("d", [0, 1, 2])
def test_kg_matches_split_form(d):
    n = (10, 10, 10)
    st, meta = air_state(10, seed=2)
    ct, s2, om = rhs(st, meta, n, scheme=O.SCHEME_KG, dirs=(d,))
    ref = kg_split_form(s2, om, d)
    scale = np.max(np.abs(ref), axis=(1, 2, 3))
    err = np.max(np.abs(ct - ref), axis=(1, 2, 3)) / scale
    assert np.all(err < 1e-12), err


def test_kg_consistency_sixth_order():
    """Single species, p const, u = 1 in x, T = p (1 + 0.2 sin x) so that
    rho = 1/(1 + 0.2 sin x): the KG mass equation is minus the 6th-order
    central derivative of rho, whose error against the analytic
    0.2 cos x/(1 + 0.2 sin x)^2 falls by ~2^6 per refinement."""
    errs = []
    for N in (16, 32):
        st, meta = tgv_state(N)
        st["velocity"] = np.zeros_like(st["velocity"])
        st["velocity"][0] = 1.0
        p0 = float(np.mean(st["pressure"]))
        st["pressure"] = np.full_like(st["pressure"], p0)
        x = cases.cell_centres(N)
        st["temperature"] = np.broadcast_to(p0 * (1.0 + 0.2 * np.sin(x))[None, None, :], st["pressure"].shape).copy()
        ct, _, _ = rhs(st, meta, (N, N, N), scheme=O.SCHEME_KG, dirs=(0,))
        exact = np.broadcast_to((0.2 * np.cos(x) / (1.0 + 0.2 * np.sin(x)) ** 2)[None, None, :], ct[0].shape)
        errs.append(float(np.max(np.abs(ct[0] - exact)) / np.max(np.abs(exact))))
    assert errs[1] < errs[0] / 40.0, errs
    assert errs[1] < 5e-5, errs


@pytest.mark.parametrize("scheme", [O.SCHEME_TENO, O.SCHEME_KG])
def test_axis_permutation(scheme):
    """The x sweep of a state equals the y sweep of the same state with x and y
    exchanged (fields transposed, u and v swapped, per-direction metrics and
    node types swapped)."""
    n = (10, 10, 10)
    st, meta = air_state(10, seed=4)
    ct, _, _ = rhs(st, meta, n, scheme=scheme, dirs=(0,))
    sw = {}
    for k, v in st.items():
        if k in sw:
            continue
        if k[-1] in "01" and k[:-1] + "0" in st and k[:-1] + "1" in st:
            a, b = k[:-1] + "0", k[:-1] + "1"
            sw[a] = st[b].swapaxes(-1, -2).copy()
            sw[b] = st[a].swapaxes(-1, -2).copy()
        else:
            sw[k] = v.swapaxes(-1, -2).copy()
    vel = sw["velocity"].copy()
    sw["velocity"][0], sw["velocity"][1] = vel[1], vel[0]
    ct2, _, _ = rhs(sw, meta, n, scheme=scheme, dirs=(1,))
    ct2 = ct2.swapaxes(-1, -2).copy()
    ns = 5
    ct2[[ns, ns + 1]] = ct2[[ns + 1, ns]]
    scale = np.max(np.abs(ct), axis=(1, 2, 3))
    err = np.max(np.abs(ct2 - ct), axis=(1, 2, 3)) / scale
    assert np.all(err < 1e-13), err


def test_diffusion_species_fluxes_sum_to_zero():
    n = (10, 10, 10)
    st, meta = air_state(10, seed=5)
    for d in range(3):
        ct, _, _ = rhs(st, meta, n, diffusion_dir=True, dirs=(d,))
        s = np.sum(ct[:5], axis=0)
        assert np.max(np.abs(s)) <= 1e-12 * np.max(np.abs(ct[:5])), d


def test_diffusion_linear_temperature_constant_properties():
    N = 12
    st, meta = tgv_state(N)
    st["velocity"] = np.zeros_like(st["velocity"])
    p0 = float(np.mean(st["pressure"]))
    st["pressure"] = np.full_like(st["pressure"], p0)
    # linear in z over the interior, evaluated away from the periodic seam
    z = np.arange(N, dtype=float)
    st["temperature"] = np.broadcast_to((100.0 + 0.5 * z)[:, None, None], st["pressure"].shape).copy()
    ct, _, _ = rhs(st, meta, (N, N, N), diffusion_dir=True, dirs=(2,))
    interior = ct[4][2:N - 2]
    assert np.max(np.abs(interior)) <= 1e-10 * np.max(np.abs(ct[4])) + 1e-300


def test_teno_all_kept_is_central_sixth_order():
    central = np.array([1.0, -8.0, 37.0, 37.0, -8.0, 1.0]) / 60.0
    rng = np.random.default_rng(0)
    for op in (0, 1, 2):
        for trial in range(20):
            a, b, c = rng.normal(size=3)
            x = np.arange(-2, 4, dtype=float)
            y = a + b * x + c * 1e-3 * np.sin(0.1 * x)
            v, bits = O.teno_reconstruct(op, 0, y)
            if all(bits):
                assert abs(v - float(np.dot(central, y))) <= 1e-13 * max(1.0, abs(v))


def test_teno_reproduces_quadratics():
    x = np.arange(-2, 4, dtype=float)
    for op in (0, 1, 2):
        for side in (0, 1):
            for (a, b, c) in ((1.0, 0.5, -0.25), (-3.0, 2.0, 0.7), (0.1, -1.0, 4.0)):
                y = a + b * x + c * x * x
                v, _ = O.teno_reconstruct(op, side, y)
                # y are cell averages of q(x) = a - c/12 + b x + c x^2; the face
                # between points 2 and 3 sits at x = 1/2 and every candidate
                # stencil reproduces q there exactly
                exact = a - c / 12.0 + 0.5 * b + 0.25 * c
                assert abs(v - exact) <= 1e-12 * max(1.0, abs(exact)), (op, side, v, exact)


def test_similarity_profile_blasius_limit():
    """The compressible similarity solution of the boundary-layer workload
    reduces to Blasius at Mach 0 with an adiabatic-free wall at the
    freestream temperature: f''(0) = 0.33206 (u/U ~ 0.33206 Y near the wall)
    and the 99 % thickness Y99 = 4.91 in units of sqrt(nu x / U)."""
    Y, u, g = cases.similarity_profile(0.0, 1.0)
    assert np.max(np.abs(g - 1.0)) < 1e-10
    slope = (u[1] - u[0]) / (Y[1] - Y[0])
    assert abs(slope - 0.33206) < 2e-4, slope
    assert abs(float(np.interp(0.99, u, Y)) - 4.91) < 0.02
