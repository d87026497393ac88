"""Closed-form pins of the oracle's characteristic-space flux reconstruction,
Roe state, sensors, viscous stress and positivity limiter (CPU only).

Each check is an identity of the mathematics that the reference algorithm
(src/prometeo_rhs.inl:200-759, src/prometeo_metric.inl:352-1038,
src/prometeo_sensor.inl:50-77, src/prometeo_rhs.inl:1058-1215) must satisfy,
chosen so that a plausible slip in the restatement -- a wrong eigenvector
entry, a sign in the Lax-Friedrichs split, a dropped term of the sound speed,
a swapped stencil row, a wrong blend -- fails it:

  TENO all kept      where every candidate stencil is kept, the TENO face
                     flux is -C6(f) of the physical flux f, whatever the Roe
                     state: -(C6(L f + lam L q) + C6(L f - lam L q))/2 back-
                     projected is -R L C6(f) = -C6(f), C6 being the 6th-order
                     interpolation [1,-8,37,37,-8,1]/60.  Pins the projections,
                     the split and the back-projection together.
  L R = I            the back-projection inverts the projection for Roe states
                     of arbitrary cell pairs.
  Roe(L, L)          equal states give the cell state back, with a^2 = gamma
                     p / rho of the ideal-gas mixture (the sound speed the
                     properties task stores).
  Ducros             a pure rotation gives exactly 0, a pure dilatation ~1.
  TENO-A sensor      constant rho Y_i is smooth everywhere; a step is flagged
                     exactly in the cells whose six-point stencil straddles it.
  TENO step          every candidate stencil straddling a step at the face is
                     dropped (only the upwind three-point stencil is kept, and
                     the face value is the upwind constant), for TENO, TENO-A
                     and TENO-LAD, both sides.
  viscous stress     u = a x^2: d(sigma_xx)/dx = (8/3) mu a; u = a y^2:
                     d(sigma_xy)/dy = 2 mu a (staggered 2nd-order differences
                     are exact on quadratics).
  limiter            a limited face flux is a convex blend theta HO + (1 -
                     theta) LO of the unlimited flux and the first-order
                     Lax-Friedrichs flux built independently from the
                     projections, and the predicted partial densities on both
                     sides are >= min(1e-60, first-order prediction).
  NSCBC              reduced states of the characteristic boundary fluxes
                     (src/prometeo_rhs.inl:1288-1631), each side of each
                     direction: a uniform state at the far-field pressure
                     leaves the RHS unchanged; a pressure offset gives the
                     relaxation wave K (p - PInf) of both K branches; a
                     normal pressure gradient gives the outgoing acoustic
                     wave; a tangential shear gives the viscous heating
                     mu (du_t/dn)^2; a mass-fraction gradient convected out
                     gives rho u dY/dn and its enthalpy term, and nothing when
                     the flow enters; a supersonic inlet clears the RHS;
                     prescribed dT/dt and du_n/dt give the inlet waves.
                     Far field: a cell at the far-field state is left alone
                     (supersonic and subsonic inflow); a temperature or
                     pressure offset from it gives the relaxation waves
                     K (TInf/T - 1) and Kp (p - PInf); an outflow at PInf
                     is left alone.  Ghost values: the NSCBC inflow ghost
                     recovers the pressure of the transported density at the
                     imposed T and X (or takes Pbc when supersonic); the
                     outflow ghost is the primitive recompute.
"""
import numpy as np
import pytest

from oracle import oracle as O
from htr_amd import cases
from hotpath import air_state, complete_state, dt_cfl, oracle_fields, oracle_mix, tgv_state

C6 = np.array([1.0, -8.0, 37.0, 37.0, -8.0, 1.0]) / 60.0


def mesh(n):
    x = cases.cell_centres(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    return tuple(a.transpose(2, 1, 0) for a in (X, Y, Z))          # (z, y, x)


def smooth_state(n, kind):
    X, Y, Z = mesh(n)
    vel = np.stack([0.3 + 0.2 * np.sin(X) * np.cos(Y), -0.1 + 0.2 * np.cos(Y + Z), 0.15 * np.sin(Z + X)])
    st = {}
    if kind == "constprop":
        _, meta = cases.tgv(n)
        st["velocity"] = vel
        st["pressure"] = 1.0 + 0.1 * np.sin(X + 2 * Y) * np.cos(Z)
        st["temperature"] = 1.0 + 0.1 * np.cos(X) * np.sin(Y + Z)
        st["MolarFracs"] = np.ones((1,) + X.shape)
    else:
        _, meta = cases.air_box(n, 0)
        Yb = np.array([0.6, 0.2, 0.1, 0.05, 0.05])[:, None, None, None]
        ph = 0.7 * np.arange(5)[:, None, None, None]
        Ym = Yb * (1 + 0.2 * np.sin(X[None] + ph) * np.cos(Y - Z)[None])
        Ym /= Ym.sum(axis=0, keepdims=True)
        W = cases.AIR_W[:, None, None, None]
        MixW = 1.0 / np.sum(Ym / W, axis=0)
        st["MolarFracs"] = Ym * MixW[None] / W
        st["velocity"] = 5.0 * vel
        st["pressure"] = 0.1 * (1.0 + 0.1 * np.sin(X + 2 * Y) * np.cos(Z))
        st["temperature"] = 8.0 * (1.0 + 0.1 * np.cos(X) * np.sin(Y + Z))
    cases.set_metrics(st, [cases.periodic_metrics(n)] * 3, (n, n, n))
    return st, meta


def all_kept_word(neq):
    """Decision word of a TENO face whose 2 nEq reconstructions kept every stencil."""
    h = 0
    for _ in range(2 * neq):
        h = (h * 17 + 15) & 0x3fffffff
    return h << 2


@pytest.mark.parametrize("kind", ["constprop", "airmix"])
@pytest.mark.parametrize("scheme", [O.SCHEME_TENO, O.SCHEME_TENOA, O.SCHEME_TENOLAD])
def test_teno_all_kept_is_minus_central6_of_physical_flux(kind, scheme):
    n = 32
    st, meta = smooth_state(n, kind)
    st, om = complete_state(st, meta, (n, n, n), sensors=False)
    ns = om.nSpec
    for d in range(3):
        g, f = oracle_fields(st, ns, (n, n, n))
        f["Conserved_t"] = 0.0
        dec = np.zeros(f["rho"].shape, np.uint32)
        O.euler_flux(g, f, d, scheme, 1.0, 0.0, 1e-6, om, dec)
        ax = 3 - d
        Q = st["Conserved"]
        u = st["velocity"][d]
        p = st["pressure"]
        F = Q * u[None]
        F[ns + d] += p
        F[ns + 3] += p * u
        Fc = sum(C6[k] * np.roll(F, -(k - 2), axis=ax) for k in range(6))
        ref = -(Fc - np.roll(Fc, 1, axis=ax)) * st["m_e%d" % d][None]
        kept = dec == all_kept_word(ns + 4)
        cells = kept & np.roll(kept, 1, axis=2 - d)
        assert cells.mean() > 0.4, cells.mean()
        err = np.max(np.abs(f["Conserved_t"] - ref) * cells[None], axis=(1, 2, 3)) / np.max(np.abs(ref), axis=(1, 2, 3))
        assert np.all(err < 1e-12), (d, err)


def cell_vector(st, ns, idx):
    z, y, x = idx
    return np.concatenate([[st["rho"][z, y, x]], st["velocity"][:, z, y, x],
                           [st["pressure"][z, y, x], st["Conserved"][ns + 3, z, y, x]], st["MassFracs"][:, z, y, x]])


@pytest.fixture(scope="module")
def air8():
    st, meta = air_state(8, seed=2)
    st, om = complete_state(st, meta, (8, 8, 8), sensors=False)
    return st, om


def test_back_projection_inverts_projection(air8):
    st, om = air8
    ns = om.nSpec
    rng = np.random.default_rng(1)
    for _ in range(30):
        a, b = tuple(rng.integers(0, 8, 3)), tuple(rng.integers(0, 8, 3))
        roe = O.roe_state(om, cell_vector(st, ns, a), cell_vector(st, ns, b))
        for d in range(3):
            q = rng.normal(size=ns + 4)
            r = O.char_project(om, d, roe, q)
            back = O.char_project(om, d, roe, r, inverse=True)
            assert np.max(np.abs(back - q)) <= 1e-12 * np.max(np.abs(q))
            # and the other order: L R = I
            again = O.char_project(om, d, roe, O.char_project(om, d, roe, q, inverse=True))
            assert np.max(np.abs(again - q)) <= 1e-12 * np.max(np.abs(q))


def test_roe_of_equal_states_is_the_cell_state(air8):
    st, om = air8
    ns = om.nSpec
    rng = np.random.default_rng(2)
    for _ in range(20):
        idx = tuple(rng.integers(0, 8, 3))
        z, y, x = idx
        v = cell_vector(st, ns, idx)
        roe = O.roe_state(om, v, v)
        rho, vel, p, E = v[0], v[1:4], v[4], v[5]
        assert roe[4] == rho and np.all(roe[5:8] == vel) and np.all(roe[9:9 + ns] == v[6:])
        assert roe[1] == E / rho - 0.5 * float(np.dot(vel, vel))
        assert roe[2] == (E + p) / rho
        # ideal-gas mixture: a^2 = gamma p / rho = the stored sound speed squared
        assert abs(roe[0] - st["SoS"][z, y, x]) <= 1e-13 * roe[0]


def test_ducros_rotation_and_dilatation():
    n = 16
    X, Y, Z = mesh(n)
    st, _ = tgv_state(n)
    rot = dict(st, velocity=np.stack([np.sin(Y), np.sin(Z), np.sin(X)]))
    g, f = oracle_fields(rot, 1, (n, n, n))
    O.ducros_sensor(g, f, 1e-3)
    assert np.all(f["DucrosSensor"] == 0.0)
    dil = dict(st, velocity=100.0 * np.stack([np.sin(X), np.sin(Y), np.sin(Z)]))
    g, f = oracle_fields(dil, 1, (n, n, n))
    O.ducros_sensor(g, f, 1e-3)
    D = f["DucrosSensor"]
    assert np.all(D > 0.99999) and np.all(D <= 1.0)


def test_tenoa_sensor_constant_and_step():
    n = 16
    st, meta = air_state(n, seed=0, composition="per-seed")
    st, om = complete_state(st, meta, (n, n, n), sensors=False)
    g, f = oracle_fields(st, 5, (n, n, n))
    rhoY = np.broadcast_to(np.array([0.5, 0.3, 0.1, 0.05, 0.05])[:, None, None, None], (5, n, n, n)).copy()
    f["Conserved"][:5] = rhoY
    f["DucrosSensor"] = 1.0
    for d in range(3):
        O.shock_sensor(g, f, d, om)
        assert np.all(f["shockSensor%d" % d] == 1), d
    step = rhoY.copy()
    step[:, :, :, n // 2:] *= 3.0
    f["Conserved"][:5] = step
    O.shock_sensor(g, f, 0, om)
    rough = np.where(f["shockSensor0"][0, 0] == 0)[0]
    # jumps between x = 7|8 and (periodically) 15|0: the sensor of cell p
    # reads p - 2 .. p + 3, which straddles the jump for p = 5 .. 9 and
    # p = 13 .. 15, 0, 1
    assert list(rough) == [0, 1, 5, 6, 7, 8, 9, 13, 14, 15]
    assert np.all(f["shockSensor0"] == f["shockSensor0"][:1, :1, :])


@pytest.mark.parametrize("op", [0, 1, 2])
@pytest.mark.parametrize("side", [0, 1])
def test_teno_drops_stencils_across_a_step(op, side):
    for jump in (1.0, -5.0, 1e3):
        y = np.array([1.0, 1.0, 1.0, 1.0 + jump, 1.0 + jump, 1.0 + jump])
        v, bits = O.teno_reconstruct(op, side, y)
        # the three-point stencil on the upwind side of the face is the only
        # one that does not straddle the step between points 2 and 3
        assert bits == [0, 0, 1, 0], (op, side, jump, bits)
        upwind = 1.0 if side == 0 else 1.0 + jump
        assert abs(v - upwind) <= 1e-14 * max(1.0, abs(upwind))


def viscous_case(n, dir_grad, comp, a=0.37):
    st, meta = tgv_state(n)
    X, Y, Z = mesh(n)
    coord = (X, Y, Z)[dir_grad]
    vel = np.zeros((3,) + X.shape)
    vel[comp] = a * coord ** 2
    st["velocity"] = vel
    p0 = float(np.mean(st["pressure"]))
    st["pressure"] = np.full_like(X, p0)
    st["temperature"] = np.full_like(X, p0)
    st, om = complete_state(st, meta, (n, n, n), sensors=False)
    g, f = oracle_fields(st, 1, (n, n, n))
    f["Conserved_t"] = 0.0
    O.diffusion_flux(g, f, dir_grad, om)
    return f["Conserved_t"], meta["visc"]


def test_viscous_stress_closed_forms():
    n, a = 16, 0.37
    ct, mu = viscous_case(n, 0, 0, a)                 # u = a x^2, sweep x
    interior = ct[1][:, :, 2:n - 2]
    assert np.max(np.abs(interior - 8.0 / 3.0 * mu * a)) <= 1e-12 * (8.0 / 3.0 * mu * a) * 1e3
    ct, mu = viscous_case(n, 1, 0, a)                 # u = a y^2, sweep y
    interior = ct[1][:, 2:n - 2, :]
    assert np.max(np.abs(interior - 2.0 * mu * a)) <= 1e-12 * (2.0 * mu * a) * 1e3
    # no normal stress from a shear flow, no mass diffusion of a single species
    assert np.max(np.abs(ct[2][:, 2:n - 2, :])) <= 1e-15 and np.max(np.abs(ct[0])) == 0.0


def test_limited_flux_is_a_positive_convex_blend():
    n = 16
    st, meta = air_state(n, seed=0)
    st, om = complete_state(st, meta, (n, n, n))
    ns = om.nSpec
    dt = dt_cfl(st, meta, (n, n, n)) * 40.0
    g, f = oracle_fields(st, ns, (n, n, n))
    rng = np.random.default_rng(5)
    for d in range(3):
        dec = np.zeros((n, n, n), np.uint32)
        O.euler_flux(g, f, d, O.SCHEME_TENO, 1.0, 0.0, dt, om, dec)
        lim = np.argwhere(dec & 2)
        assert len(lim) > 0
        for (z, y, x) in lim[rng.choice(len(lim), size=min(40, len(lim)), replace=False)]:
            F, w = O.face_flux(g, f, d, 0, (int(x), int(y), int(z)), 1.0, 0.0, dt, om)
            HO, w0 = O.face_flux(g, f, d, 0, (int(x), int(y), int(z)), 1.0, 0.0, 1e-300, om)
            assert (w & 2) and not (w0 & 2)
            p = (z, y, x)
            pp = list(p)
            pp[2 - d] = (pp[2 - d] + 1) % n
            pp = tuple(pp)
            roe = O.roe_state(om, cell_vector(st, ns, p), cell_vector(st, ns, pp))
            pts = []
            for k in (-2, -1, 0, 1, 2, 3):
                q = list(p)
                q[2 - d] = (q[2 - d] + k) % n
                pts.append(tuple(q))
            un = np.array([st["velocity"][d][q] for q in pts])
            c = np.array([st["SoS"][q] for q in pts])
            lam = np.full(ns + 4, np.max(np.abs(un)))
            lam[0], lam[-1] = np.max(np.abs(un - c)), np.max(np.abs(un + c))

            def phys(q):
                Q = st["Conserved"][(slice(None),) + q]
                u, pr = st["velocity"][d][q], st["pressure"][q]
                fl = Q * u
                fl[ns + d] += pr
                fl[ns + 3] += pr * u
                return Q, fl
            Qa, Fa = phys(p)
            Qb, Fb = phys(pp)
            Wp = O.char_project(om, d, roe, Fa) + lam * O.char_project(om, d, roe, Qa)
            Wm = O.char_project(om, d, roe, Fb) - lam * O.char_project(om, d, roe, Qb)
            LO = O.char_project(om, d, roe, -0.5 * (Wp + Wm), inverse=True)
            k = int(np.argmax(np.abs(HO - LO)))
            theta = (F[k] - LO[k]) / (HO[k] - LO[k])
            assert -1e-12 <= theta <= 1.0 + 1e-12
            assert np.max(np.abs(F - (theta * HO + (1 - theta) * LO))) <= 1e-13 * np.max(np.abs(F))
            lim_f = dt * (st["m_e%d" % d][p] + st["m_e%d" % d][pp])
            rp = st["Conserved"][:ns][(slice(None),) + p]
            rq = st["Conserved"][:ns][(slice(None),) + pp]
            pred = np.concatenate([rp + lim_f * F[:ns], rq - lim_f * F[:ns]])
            pred_lo = np.concatenate([rp + lim_f * LO[:ns], rq - lim_f * LO[:ns]])
            assert np.all(pred >= np.minimum(1e-60, pred_lo) - 1e-15 * np.max(np.abs(rp)))


@pytest.mark.parametrize("side", range(6))
def test_bc_ghost_values(side):
    """Ghost cells of one boundary (src/prometeo_bc.rg:1159-1265): the wall
    tasks take X and p (and T for the adiabatic wall) from the interior
    neighbour one cell inward along the normal and set u = 0; the isothermal
    wall takes T from the profile; the Dirichlet task takes X, u, T from the
    profiles and p from its argument.  The two cells on either side of the
    ghost layer hold different values, so a neighbour on the wrong side or a
    swapped field fails."""
    n = (6, 7, 5)
    rng = np.random.default_rng(side)
    st, meta = air_state(8, seed=1)
    om = oracle_mix(meta)
    g = O.Grid(*n)
    f = O.Fields(g, 5)
    for k, nc in (("MolarFracs", 5), ("velocity", 3), ("MolarFracs_profile", 5), ("velocity_profile", 3)):
        f[k] = rng.uniform(0.1, 1.0, size=(nc,) + f["rho"].shape)
    for k in ("pressure", "temperature", "temperature_profile"):
        f[k] = rng.uniform(0.5, 2.0, size=f["rho"].shape)
    d, neg = side // 2, side % 2 == 0
    lo, hi = [0, 0, 0], [n[0] - 1, n[1] - 1, n[2] - 1]
    lo[d] = hi[d] = 0 if neg else n[d] - 1
    ax = 2 - d                                   # array axis of direction d in (z, y, x)
    gi = 0 if neg else n[d] - 1
    ni = 1 if neg else n[d] - 2
    take = lambda a, i: np.take(a, i, axis=a.ndim - 3 + ax)
    before = {k: f[k].copy() for k in f.a}
    for kind in (0, 1, 2):
        for k, v in before.items():
            f[k] = v
        O.set_bc(g.with_rect(tuple(lo), tuple(hi)), f, kind, side, 0.77, om)
        if kind == 0:
            assert np.all(take(f["MolarFracs"], gi) == take(before["MolarFracs_profile"], gi))
            assert np.all(take(f["velocity"], gi) == take(before["velocity_profile"], gi))
            assert np.all(take(f["temperature"], gi) == take(before["temperature_profile"], gi))
            assert np.all(take(f["pressure"], gi) == 0.77)
        else:
            assert np.all(take(f["MolarFracs"], gi) == take(before["MolarFracs"], ni))
            assert np.all(take(f["pressure"], gi) == take(before["pressure"], ni))
            assert np.all(take(f["velocity"], gi) == 0.0)
            Tsrc = take(before["temperature"], ni) if kind == 1 else take(before["temperature_profile"], gi)
            assert np.all(take(f["temperature"], gi) == Tsrc)
        # nothing outside the ghost layer changes
        for k in ("MolarFracs", "pressure", "temperature", "velocity"):
            a, b = f[k], before[k]
            keep = [i for i in range(n[d]) if i != gi]
            assert np.all(np.take(a, keep, axis=a.ndim - 3 + ax) == np.take(b, keep, axis=b.ndim - 3 + ax))


# ---------------------------------------------------------------- NSCBC
NSCBC_N = (6, 7, 5)


def nscbc_case(side, seed=0):
    """Air fields on a small box with collocated boundary node types along the
    normal of the side (L_C .. R_C, Grad rows (0, 1) and (1, 0) at the ends)
    and a uniform reference state; returns (grid on the boundary layer of
    the side, fields, mixture, helpers)."""
    n = NSCBC_N
    d, minus = side // 2, side % 2 == 0
    st, meta = air_state(8, seed=1)
    om = oracle_mix(meta)
    g = O.Grid(*n)
    f = O.Fields(g, 5)
    rng = np.random.default_rng(seed + 17 * side)
    Y = np.array([0.70, 0.20, 0.05, 0.03, 0.02])
    f["MassFracs"] = Y[:, None, None, None]
    f["rho"], f["SoS"], f["temperature"], f["pressure"], f["mu"] = 1.3, 1.7, 1.1, 0.9, 2e-3
    f["velocity"] = 0.0
    f["Conserved"] = rng.uniform(0.5, 1.5, size=f["Conserved"].shape)
    f["Conserved_t"] = rng.uniform(-1.0, 1.0, size=f["Conserved_t"].shape)
    nt = np.full(n[d], cases.STD, np.int32)
    nt[0], nt[1] = cases.L_C, cases.LP1_C
    nt[-1], nt[-2], nt[-3] = cases.R_C, cases.RM1_C, cases.RM2_C
    shape = [1, 1, 1]
    shape[2 - d] = n[d]
    f["nType%d" % d] = nt.reshape(shape)
    for dd in range(3):
        f["m_d%d" % dd] = 2.5 if dd == d else 1.3
    lo, hi = [0, 0, 0], [n[0] - 1, n[1] - 1, n[2] - 1]
    lo[d] = hi[d] = 0 if minus else n[d] - 1
    ax = 2 - d
    bi, ii = (0, 1) if minus else (n[d] - 1, n[d] - 2)

    def bnd(a):
        return np.take(a, bi, axis=a.ndim - 3 + ax)

    def along(x):
        """a field linear along the normal: x * index"""
        sh = [1, 1, 1]
        sh[ax] = n[d]
        return (np.arange(n[d], dtype=np.float64) * x).reshape(sh)

    return g.with_rect(tuple(lo), tuple(hi)), f, om, dict(d=d, minus=minus, bnd=bnd, along=along, ii=ii, bi=bi)


def nscbc_expect(f0, h, ns, d1, dM, rhoLS=None, dE_extra=0.0, dN=0.0, cpT=0.0):
    """Conserved_t at the boundary cells after the LODI update with the given
    wave terms (the reduced form of the outflow update)."""
    d = h["d"]
    bnd = h["bnd"]
    Ct = bnd(f0["Conserved_t"]).copy()
    Y = bnd(f0["MassFracs"])
    rho, P = bnd(f0["rho"]), bnd(f0["pressure"])
    u = bnd(f0["velocity"])
    C = bnd(f0["Conserved"])
    for s in range(ns):
        Ct[s] -= d1 * Y[s] + (rhoLS[s] if rhoLS is not None else 0.0)
    for c in range(3):
        Ct[ns + c] -= u[c] * d1 + rho * dM[c]
    Ct[ns + 3] -= (C[ns + 3] + P) * d1 / rho + sum(C[ns + c] * dM[c] for c in range(3)) + cpT * dN - dE_extra
    return Ct


@pytest.mark.parametrize("side", range(6))
def test_nscbc_outflow_closed_forms(side):
    g, f, om, h = nscbc_case(side)
    d, minus, bnd = h["d"], h["minus"], h["bnd"]
    rho, c, P = 1.3, 1.7, 0.9
    ns = 5
    f0 = f.copy()
    # (a) uniform at rest, p = PInf: unchanged
    O.nscbc_outflow(g, f, side, 0.3, 2.0, P, om)
    assert np.array_equal(f["Conserved_t"], f0["Conserved_t"])
    # (b) pressure offset, both K branches; the relaxed wave is LN on a Minus
    # side (incoming from the far field), L1 on a Plus side
    for Mmax in (0.3, 1.2):
        f = f0.copy()
        un = -0.2 if minus else 0.2                # leaving the domain, subsonic
        f["velocity"][d] = un
        PInf, ell = 0.6, 2.0
        O.nscbc_outflow(g, f, side, Mmax, ell, PInf, om)
        K = 0.25 * (1 - Mmax * Mmax) * c / ell if Mmax < 0.99 else 0.25 * (c - un * un / c) / ell
        Lr = K * (P - PInf)
        d1 = 0.5 * Lr / c**2
        dM = [0.0, 0.0, 0.0]
        dM[d] = (Lr / (2 * rho * c)) if minus else (-Lr / (2 * rho * c))
        f0b = f0.copy()
        f0b["velocity"][d] = un
        want = nscbc_expect(f0b, h, ns, d1, dM)
        np.testing.assert_allclose(bnd(f["Conserved_t"]), want, rtol=1e-13, atol=1e-15)
    # (c) normal pressure gradient at rest, p = PInf on the boundary: the
    # outgoing acoustic wave (L1 = -c dp/dn on a Minus side, LN = c dp/dn on a Plus side)
    f = f0.copy()
    f["pressure"] = P + (h["along"](0.01) - h["along"](0.01).take([h["bi"]], axis=2 - d))
    dp_dn = 0.01 * 2.5          # slope along the index direction, times m_d
    O.nscbc_outflow(g, f, side, 0.3, 2.0, P, om)
    if minus:
        L1, LN = -c * dp_dn, 0.0
    else:
        L1, LN = 0.0, c * dp_dn
    d1 = 0.5 * (L1 + LN) / c**2
    dM = [0.0, 0.0, 0.0]
    dM[d] = (LN - L1) / (2 * rho * c)
    want = nscbc_expect(f0, h, ns, d1, dM)
    np.testing.assert_allclose(bnd(f["Conserved_t"]), want, rtol=1e-12, atol=1e-15)
    # (d) tangential shear at rest: viscous heating mu (du_t/dn)^2 only
    t1 = 1 if d == 0 else 0
    for t in (t1, 3 - d - t1):
        f = f0.copy()
        f["velocity"][t] = h["along"](0.03)
        O.nscbc_outflow(g, f, side, 0.3, 2.0, P, om)
        dudn = 0.03 * 2.5
        want = bnd(f0["Conserved_t"]).copy()
        want[ns + 3] += 2e-3 * dudn * dudn
        np.testing.assert_allclose(bnd(f["Conserved_t"]), want, rtol=1e-12, atol=1e-15)
    # (f) normal velocity linear along the normal, at rest on the boundary:
    # the outgoing acoustic wave of -rho c^2 du/dn, tau_nn = 4/3 mu du/dn the
    # same at the boundary and the interior point (d tau/dn = 0), viscous
    # heating d(u tau)/dn = tau du/dn
    f = f0.copy()
    a = 0.05
    f["velocity"][d] = h["along"](a) - h["along"](a).take([h["bi"]], axis=2 - d)
    O.nscbc_outflow(g, f, side, 0.3, 2.0, P, om)
    dudn = a * 2.5
    Lw = rho * c * c * dudn                        # L1 on a Minus side, LN on a Plus side
    d1 = 0.5 * Lw / c**2
    dM = [0.0, 0.0, 0.0]
    dM[d] = (-Lw if minus else Lw) / (2 * rho * c)
    tau = 2e-3 * 4 * dudn / 3
    want = nscbc_expect(f0, h, ns, d1, dM, dE_extra=tau * dudn)
    np.testing.assert_allclose(bnd(f["Conserved_t"]), want, rtol=1e-12, atol=1e-15)
    # (e) a mass-fraction gradient convected out of the domain: rho u dY/dn
    # and its enthalpy term; flow entering the domain leaves the species alone
    for leaving in (True, False):
        f = f0.copy()
        un = (-0.3 if minus else 0.3) * (1 if leaving else -1)
        f["velocity"][d] = un
        dY = np.array([0.01, -0.004, -0.003, -0.002, -0.001])
        for s in range(ns):
            f["MassFracs"][s] = f0["MassFracs"][s] + h["along"](dY[s]) - h["along"](dY[s]).take([h["bi"]], axis=2 - d)
        O.nscbc_outflow(g, f, side, 0.3, 2.0, P, om)
        want = bnd(f0["Conserved_t"]).copy()
        if leaving:
            T = 1.1
            Yb = np.array([0.70, 0.20, 0.05, 0.03, 0.02])
            cpT = O.call("o_GetHeatCapacity", om, T, Yb) * T
            MixW = O.molar_weight_from_Y(om, Yb)
            for s in range(ns):
                LS = un * dY[s] * 2.5
                W = 1.0 / (1.0 / O.molar_weight_from_Y(om, np.eye(ns)[s]))
                hs = O.call("o_GetSpeciesEnthalpy", om, s, T)
                want[s] -= rho * LS
                want[ns + 3] -= rho * LS * (hs - cpT * MixW / W)
        np.testing.assert_allclose(bnd(f["Conserved_t"]), want, rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("side", range(6))
def test_nscbc_inflow_closed_forms(side):
    g, f, om, h = nscbc_case(side)
    d, minus, bnd = h["d"], h["minus"], h["bnd"]
    rho, c, T = 1.3, 1.7, 1.1
    ns = 5
    Yb = np.array([0.70, 0.20, 0.05, 0.03, 0.02])
    f0 = f.copy()
    # supersonic inlet: the RHS of the boundary cells is cleared
    f["velocity"][d] = 1.9 if minus else -1.9
    O.nscbc_inflow(g, f, side, om)
    assert np.all(bnd(f["Conserved_t"]) == 0.0)
    # subsonic, uniform, prescribed dT/dt: rho Y dT/dt / T off the species only
    f = f0.copy()
    f["velocity"][d] = 0.4 if minus else -0.4
    f["dTdtBoundary"] = 0.37
    O.nscbc_inflow(g, f, side, om)
    want = bnd(f0["Conserved_t"]).copy()
    for s in range(ns):
        want[s] -= (rho * 0.37 / T) * Yb[s]
    np.testing.assert_allclose(bnd(f["Conserved_t"]), want, rtol=1e-13, atol=1e-15)
    # subsonic, uniform, prescribed du_n/dt = b: d1 = -/+ (rho b / c + rho c b / (cp T))
    f = f0.copy()
    f["velocity"][d] = 0.4 if minus else -0.4
    b = 0.21
    f["dudtBoundary"][d] = b
    O.nscbc_inflow(g, f, side, om)
    cp = O.call("o_GetHeatCapacity", om, T, Yb)
    d1 = (rho * b / c + rho * c * b / (cp * T)) * (-1 if minus else 1)
    want = bnd(f0["Conserved_t"]).copy()
    for s in range(ns):
        want[s] -= d1 * Yb[s]
    np.testing.assert_allclose(bnd(f["Conserved_t"]), want, rtol=1e-12, atol=1e-15)
    # the momentum and energy rows are never touched by the subsonic inlet
    assert np.array_equal(bnd(f["Conserved_t"])[ns:], bnd(f0["Conserved_t"])[ns:])


def test_max_mach():
    g, f, om, h = nscbc_case(0)
    gg = O.Grid(*NSCBC_N)
    f["velocity"][1] = 0.1
    f["velocity"][1][2, 3, 4] = -1.02
    assert O.max_mach(gg, f, 1) == 1.02 / 1.7
    assert O.max_mach(gg, f, 0) == 0.0


def far_field_state(f, om, PInf, TInf, Xinf, vInf):
    """Fills the cells of f with the far-field state itself (primitive,
    properties used by the far-field fluxes, conserved)."""
    MixW = O.molar_weight_from_X(om, Xinf)
    Y = O.mass_fractions(om, MixW, Xinf)
    rho = O.call("o_GetRho", om, PInf, TInf, MixW)
    e = O.call("o_GetInternalEnergy", om, TInf, Y)
    f["MassFracs"] = Y[:, None, None, None]
    f["MolarFracs_profile"] = np.asarray(Xinf)[:, None, None, None]
    f["temperature_profile"] = TInf
    f["velocity_profile"] = np.asarray(vInf)[:, None, None, None]
    f["temperature"], f["pressure"], f["rho"] = TInf, PInf, rho
    f["velocity"] = np.asarray(vInf)[:, None, None, None]
    C = np.concatenate([rho * Y, rho * np.asarray(vInf), [rho * (e + 0.5 * float(np.dot(vInf, vInf)))]])
    f["Conserved"] = C[:, None, None, None]
    return rho, Y


@pytest.mark.parametrize("side", range(6))
def test_nscbc_far_field_closed_forms(side):
    g, f, om, h = nscbc_case(side)
    d, minus, bnd = h["d"], h["minus"], h["bnd"]
    ns, c, ell = 5, 1.7, 0.9
    PInf, TInf = 0.9, 1.1
    Xinf = np.array([0.75, 0.2, 0.03, 0.01, 0.01])
    sgn = 1 if minus else -1                       # + normal velocity enters on a Minus side
    for Ma in (1.4, 0.35):                         # supersonic, subsonic inflow
        vInf = np.array([0.1, -0.05, 0.07])
        vInf[d] = sgn * Ma * c
        rho, Y = far_field_state(f, om, PInf, TInf, Xinf, vInf)
        f0 = f.copy()
        O.nscbc_farfield(g, f, side, PInf, 0.3, ell, om)
        scale = np.abs(bnd(f0["Conserved"])).max(axis=(1, 2))[:, None, None] * c / ell
        err = np.abs(bnd(f["Conserved_t"]) - bnd(f0["Conserved_t"])) / scale
        assert err.max() < 1e-13, (Ma, err.max())
    # subsonic inflow, T off the far-field value: d1 = rho K (TInf/T - 1),
    # dN = -d1, K = 0.25 * 40 Ma^2 c / L
    Ma = 0.35
    vInf = np.array([0.1, -0.05, 0.07])
    vInf[d] = sgn * Ma * c
    rho, Y = far_field_state(f, om, PInf, TInf, Xinf, vInf)
    T = 1.25
    f["temperature"] = T
    f0 = f.copy()
    O.nscbc_farfield(g, f, side, PInf, 0.3, ell, om)
    K = 0.25 * 40 * Ma * Ma * c / ell
    d1 = rho * K * (TInf / T - 1)
    cpT = O.call("o_GetHeatCapacity", om, T, Y) * T
    want = nscbc_expect(f0, h, ns, d1, [0.0, 0.0, 0.0], dN=-d1, cpT=cpT)
    np.testing.assert_allclose(bnd(f["Conserved_t"]), want, rtol=1e-11, atol=1e-12)
    # subsonic inflow, p off PInf: the relaxed wave is LN on a Minus side, L1
    # on a Plus side, Kp (p - PInf) with Kp = 0.25 (1 + 39 Ma^2) c / L
    rho, Y = far_field_state(f, om, PInf, TInf, Xinf, vInf)
    P = 1.05
    f["pressure"] = P
    f0 = f.copy()
    O.nscbc_farfield(g, f, side, PInf, 0.3, ell, om)
    Kp = 0.25 * (1 + 39 * Ma * Ma) * c / ell
    Lr = Kp * (P - PInf)
    cpT = O.call("o_GetHeatCapacity", om, TInf, Y) * TInf
    L2 = -rho * c * c * (Lr / (2 * rho * cpT))
    d1 = (0.5 * Lr - L2) / c**2
    dM = [0.0, 0.0, 0.0]
    dM[d] = (Lr if minus else -Lr) / (2 * rho * c)
    want = nscbc_expect(f0, h, ns, d1, dM, dN=L2 / c**2, cpT=cpT)
    np.testing.assert_allclose(bnd(f["Conserved_t"]), want, rtol=1e-11, atol=1e-12)
    # outflow at PInf, uniform: left alone; with p off PInf: relaxation as in
    # the outflow task (no viscous terms)
    vInf[d] = -sgn * 0.3 * c
    rho, Y = far_field_state(f, om, PInf, TInf, Xinf, vInf)
    f0 = f.copy()
    O.nscbc_farfield(g, f, side, PInf, 0.3, ell, om)
    assert np.array_equal(f["Conserved_t"], f0["Conserved_t"])
    f["pressure"] = P
    f0 = f.copy()
    O.nscbc_farfield(g, f, side, PInf, 0.3, ell, om)
    Lr = 0.25 * (1 - 0.09) * c / ell * (P - PInf)
    dM = [0.0, 0.0, 0.0]
    dM[d] = (Lr if minus else -Lr) / (2 * rho * c)
    want = nscbc_expect(f0, h, ns, 0.5 * Lr / c**2, dM)
    np.testing.assert_allclose(bnd(f["Conserved_t"]), want, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("side", [0, 3, 5])
def test_nscbc_ghost_values(side):
    """SetNSCBC_InflowBC: X, T, u from the profiles; p from the transported
    density at the imposed T and X (the pressure the density was made from),
    Pbc where the imposed normal velocity is supersonic.  SetNSCBC_OutflowBC:
    the primitives of the transported conserved state."""
    g, f, om, h = nscbc_case(side)
    d, bnd = h["d"], h["bnd"]
    PInf, TInf, c = 0.9, 1.1, 1.7
    Xinf = np.array([0.75, 0.2, 0.03, 0.01, 0.01])
    vInf = np.array([0.3, -0.2, 0.1])
    far_field_state(f, om, PInf, TInf, Xinf, vInf)
    f["SoS"] = c
    f["pressure"] = 0.0
    O.set_bc(g, f, 3, side, 7.0, om)
    np.testing.assert_allclose(bnd(f["pressure"]), PInf, rtol=1e-14)
    assert np.all(bnd(f["temperature"]) == TInf)
    f["velocity_profile"][d] = 1.01 * c
    O.set_bc(g, f, 3, side, 7.0, om)
    assert np.all(bnd(f["pressure"]) == 7.0)
    assert np.all(bnd(f["velocity"])[d] == 1.01 * c)
    # outflow ghost: primitives of the conserved state (temperature guess off)
    far_field_state(f, om, PInf, TInf, Xinf, vInf)
    f["temperature"] = 1.3
    f["pressure"] = 0.0
    f["velocity"] = 0.0
    O.set_bc(g, f, 4, side, 0.0, om)
    np.testing.assert_allclose(bnd(f["pressure"]), PInf, rtol=1e-10)
    np.testing.assert_allclose(bnd(f["temperature"]), TInf, rtol=1e-10)
    np.testing.assert_allclose(bnd(f["velocity"]), np.broadcast_to(vInf[:, None, None], bnd(f["velocity"]).shape),
                               rtol=1e-13)
