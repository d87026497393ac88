"""Workload generators for the hot path (numpy, host side).

Each generator returns (state, meta): state maps field names to float64
numpy arrays in the FieldSet layout ((ncomp, nz, ny, nx) or (nz, ny, nx));
meta holds the mixture parameters and scalars of the case.

  tgv(n)             3-D periodic Taylor-Green vortex, calorically perfect
                     gas (gamma 1.4, R 1), Mach 0.1, Re 1600 on [0, 2 pi]^3
                     (src/prometeo_init.rg:134-158; solverTests/3DPeriodic)
  air_box(n, seed)   periodic 5-species air box: Dirichlet(8,2,.5,.5,.5) mass
                     fractions per cell, T = T0 (1 + 0.05 xi) with
                     T0 ~ U(2500, 7000) K per seed, p = 1e4 Pa, random-phase
                     solenoidal velocity with E(k) ~ k^-5/3 (1 <= k <= kmax)
                     at Mach_rms 0.4
  boundary_layer(nx, ny, nz, seed)
                     flat-plate boundary layer, Mach 6 air at 220 K over an
                     isothermal 300 K wall: x periodic (nx), y wall-normal with
                     staggered wall ghosts and tanh stretching (max adjacent
                     spacing ratio 1.04), z periodic; the laminar compressible
                     similarity profile (Chapman-Rubesin parameter 1) plus 1 %
                     seeded white noise on the velocity
  chem30_box(n, seed) cells of the air box (same seed: p, u, and T mapped
                     linearly from [2500, 7000] K onto [1500, 3000] K, inside
                     the validity range of the 30-species thermodynamic data)
                     with 30-species mass fractions ~ Dirichlet(1_30) per cell
                     (seeded), for the synthetic mechanism of
                     htr_amd.mechanism.synthetic_30
  wall_metrics(...)  node types and metrics of a direction bounded by
                     staggered walls with tanh stretching
                     (src/prometeo_grid.rg:139-200, src/prometeo_metric.cc:66-100)
"""
import math

import numpy as np

RGAS = 8.3144598
AIR_W = np.array([2 * 14.0067e-3, 2 * 15.9994e-3, 14.0067e-3 + 15.9994e-3, 14.0067e-3, 15.9994e-3])


def periodic_metrics(n, width=2 * math.pi):
    """Uniform periodic direction: every node standard, all metrics 1/dx."""
    dx = width / n
    return np.zeros(n, np.int32), np.full(n, 1.0 / dx), np.full(n, 1.0 / dx), np.full(n, 1.0 / dx)


def set_metrics(state, dims_metrics, n):
    """Broadcasts per-direction 1-D metric profiles to 3-D fields."""
    nx, ny, nz = n
    for d, (nt, me, md, ms) in enumerate(dims_metrics):
        shape = [1, 1, 1]
        shape[2 - d] = -1
        state["nType%d" % d] = np.broadcast_to(nt.reshape(shape), (nz, ny, nx)).astype(np.int32).copy()
        state["m_e%d" % d] = np.broadcast_to(me.reshape(shape), (nz, ny, nx)).copy()
        state["m_d%d" % d] = np.broadcast_to(md.reshape(shape), (nz, ny, nx)).copy()
        state["m_s%d" % d] = np.broadcast_to(ms.reshape(shape), (nz, ny, nx)).copy()


def cell_centres(n, width=2 * math.pi):
    return (np.arange(n) + 0.5) * width / n


def tgv(n=32, mach=0.1, reynolds=1600.0, gamma=1.4):
    nx = ny = nz = n
    x = cell_centres(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")          # (x, y, z)
    X, Y, Z = (a.transpose(2, 1, 0) for a in (X, Y, Z))    # -> (z, y, x)
    rho0, V0 = 1.0, 1.0
    p0 = rho0 * V0 ** 2 / (gamma * mach ** 2)
    state = {}
    state["velocity"] = np.stack([V0 * np.sin(X) * np.cos(Y) * np.cos(Z),
                                  -V0 * np.cos(X) * np.sin(Y) * np.cos(Z),
                                  np.zeros_like(X)])
    p = p0 + rho0 * V0 ** 2 / 16.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2)
    state["pressure"] = p
    state["temperature"] = p / (1.0 * rho0)
    state["MolarFracs"] = np.ones((1, nz, ny, nx))
    set_metrics(state, [periodic_metrics(n)] * 3, (nx, ny, nz))
    meta = dict(kind="constprop", R=1.0, gamma=gamma, prandtl=0.71, visc=rho0 * V0 * 1.0 / reynolds,
                dt=5.0e-3, vorticityScale=1.0)
    return state, meta


def solenoidal_velocity(n, rng, kmax, urms):
    """Random-phase divergence-free field with E(k) ~ k^-5/3 on 1 <= |k| <= kmax."""
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    KX, KY, KZ = np.meshgrid(k1, k1, k1, indexing="ij")
    K = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    shell = (K >= 1) & (K <= kmax)
    amp = np.zeros_like(K)
    amp[shell] = K[shell] ** (-5.0 / 3.0 / 2.0 - 1.0)       # sqrt(E(k)/(4 pi k^2))
    u = []
    for _ in range(3):
        ph = rng.uniform(0, 2 * np.pi, K.shape)
        u.append(amp * np.exp(1j * ph))
    # project out the dilatational part
    with np.errstate(invalid="ignore", divide="ignore"):
        kdotu = (KX * u[0] + KY * u[1] + KZ * u[2]) / np.where(K > 0, K ** 2, 1)
    u = [u[0] - KX * kdotu, u[1] - KY * kdotu, u[2] - KZ * kdotu]
    v = np.stack([np.real(np.fft.ifftn(c)) for c in u])     # (3, x, y, z)
    v = v.transpose(0, 3, 2, 1).copy()                      # -> (3, z, y, x)
    rms = math.sqrt(np.mean(np.sum(v ** 2, axis=0)))
    return v * (urms / rms)


AIR_REFS = dict(PRef=1.0e5, TRef=500.0, LRef=1.0, xref={"N2": 0.79, "O2": 0.21})


def air_refs():
    xr = AIR_REFS["xref"]
    MixWRef = xr["N2"] * AIR_W[0] + xr["O2"] * AIR_W[1]
    rhoRef = AIR_REFS["PRef"] * MixWRef / (RGAS * AIR_REFS["TRef"])
    return MixWRef, rhoRef, math.sqrt(AIR_REFS["PRef"] / rhoRef)


def air_box(n=32, seed=0, p=1.0e4, mach_rms=0.4, kmax=None, T_noise=0.05, sound_speed=None,
            composition="per-seed"):
    """Returns the primitive state (p, T, X, u) of the periodic 5-species box;
    properties and conserved variables come from the UpdateProperties /
    UpdateConserved tasks.  sound_speed(T0, Ybar) -> m/s sets the velocity
    scale (defaults to an ideal-gas estimate with gamma 1.3).

    composition "per-seed": one Dirichlet draw of the mass fractions per seed
    (uniform composition; the grid-scale variation comes from the 5 %
    temperature noise and the velocity field: about 20 % of the faces take the
    TENO branch of the hybrid scheme); "per-cell": an independent draw in every
    cell (white-noise species fields: about 85 % TENO faces)."""
    rng = np.random.default_rng(seed)
    nx = ny = nz = n
    alpha = np.array([8.0, 2.0, 0.5, 0.5, 0.5])
    if composition == "per-cell":
        Y = rng.dirichlet(alpha, size=(nz, ny, nx)).transpose(3, 0, 1, 2).copy()
    else:
        Y = np.broadcast_to(rng.dirichlet(alpha)[:, None, None, None], (5, nz, ny, nx)).copy()
    Y = np.maximum(Y, 1e-12)
    Y /= Y.sum(axis=0, keepdims=True)
    T0 = rng.uniform(2500.0, 7000.0)
    T = np.clip(T0 * (1.0 + T_noise * rng.standard_normal((nz, ny, nx))), 250.0, 19000.0)
    MixW = 1.0 / np.sum(Y / AIR_W[:, None, None, None], axis=0)
    X = Y * MixW[None] / AIR_W[:, None, None, None]
    Ybar = Y.reshape(5, -1).mean(axis=1)
    if sound_speed is None:
        Wbar = 1.0 / np.sum(Ybar / AIR_W)
        c0 = math.sqrt(1.3 * RGAS / Wbar * T0)
    else:
        c0 = sound_speed(T0, Ybar)
    kmax = kmax if kmax is not None else max(2, n // 4)
    u = solenoidal_velocity(n, rng, kmax, mach_rms * c0)     # m/s
    _, _, uRef = air_refs()
    state = {"pressure": np.full((nz, ny, nx), p / AIR_REFS["PRef"]),
             "temperature": T / AIR_REFS["TRef"],
             "MolarFracs": X, "velocity": u / uRef}
    set_metrics(state, [periodic_metrics(n)] * 3, (nx, ny, nz))
    meta = dict(kind="airmix", T0=T0, c0=c0, vorticityScale=1.0, composition=composition, **AIR_REFS)
    return state, meta


def chem30_box(n=16, seed=0, mech_seed=0):
    from . import mechanism
    st, meta = air_box(n, seed)
    mech = mechanism.synthetic_30(mech_seed)
    rng = np.random.default_rng(seed + 1000)
    nz = ny = nx = n
    Y = rng.dirichlet(np.ones(30), size=(nz, ny, nx)).transpose(3, 0, 1, 2).copy()
    W = np.asarray(mech["W"])[:, None, None, None]
    MixW = 1.0 / np.sum(Y / W, axis=0)
    T = st["temperature"] * AIR_REFS["TRef"]
    Tc = 1500.0 + (T - 2500.0) / 4500.0 * 1500.0
    state = {"pressure": st["pressure"], "temperature": Tc / mech["TRef"],
             "MolarFracs": Y * MixW[None] / W, "velocity": st["velocity"]}
    for k in st:
        if k.startswith(("nType", "m_")):
            state[k] = st[k]
    meta = dict(kind="mechanism", mech=mech, seed=seed, vorticityScale=1.0)
    return state, meta


def stretching_for_ratio(n, ratio, lo=1e-3, hi=6.0):
    """tanh stretching parameter whose largest adjacent-spacing ratio of the
    n interior cells is `ratio` (bisection)."""
    def worst(s):
        x = tanh_minus_coords(n, 1.0, s)[1:-1]
        dx = np.diff(x)
        r = dx[1:] / dx[:-1]
        return float(np.max(np.maximum(r, 1.0 / r)))
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        if worst(mid) > ratio:
            hi = mid
        else:
            lo = mid
    return lo


def similarity_profile(mach, Tw_over_Tinf, prandtl=0.72, gamma=1.4, eta_max=14.0, n=2801):
    """Laminar compressible flat-plate boundary layer (similarity solution with
    Chapman-Rubesin parameter 1, rho mu = const): in the Howarth-Dorodnitsyn
    variable eta,
        f''' + f f''/2 = 0,   g'' + Pr f g'/2 + Pr (gamma - 1) M^2 f''^2 = 0,
        f(0) = f'(0) = 0, f'(inf) = 1,  g(0) = Tw/Tinf, g(inf) = 1,
    with u/U = f' and T/Tinf = g.  The wall-normal distance in units of
    sqrt(nu_inf x / U) is Y(eta) = integral of g from 0 to eta (rho ~ 1/T at
    constant pressure).  Returns (Y, u/U, T/Tinf) on n points."""
    from scipy.integrate import solve_bvp, cumulative_trapezoid
    eta = np.linspace(0.0, eta_max, n)

    def rhs(x, y):
        f, fp, fpp, g, gp = y
        return np.vstack([fp, fpp, -0.5 * f * fpp, gp,
                          -0.5 * prandtl * f * gp - prandtl * (gamma - 1.0) * mach ** 2 * fpp ** 2])

    def bc(a, b):
        return np.array([a[0], a[1], b[1] - 1.0, a[3] - Tw_over_Tinf, b[3] - 1.0])

    guess = np.zeros((5, n))
    guess[1] = np.tanh(eta / 2.0)
    guess[0] = 2.0 * np.log(np.cosh(eta / 2.0))
    guess[2] = 0.5 / np.cosh(eta / 2.0) ** 2
    guess[3] = 1.0 + (Tw_over_Tinf - 1.0) * np.exp(-eta / 2.0) + 0.5 * (gamma - 1.0) * mach ** 2 * \
        np.exp(-eta ** 2 / 8.0)
    sol = solve_bvp(rhs, bc, eta, guess, tol=1e-8, max_nodes=100000)
    if not sol.success:
        raise RuntimeError("similarity solution did not converge: " + sol.message)
    y = sol.sol(eta)
    Y = cumulative_trapezoid(y[3], eta, initial=0.0)
    return Y, y[1], y[3]


def boundary_layer(nx=512, ny=256, nz=128, seed=0, mach=6.0, T_inf=220.0, T_wall=300.0, p=1.0e4,
                   Lx=0.5, Ly=0.05, Lz=0.1, delta=0.01, noise=0.01):
    """Flat-plate boundary layer of 5-species air: Mach `mach` freestream at
    T_inf over an isothermal wall at T_wall (y = 0, staggered ghost below the
    first cell), wall-normal tanh stretching with largest adjacent-spacing
    ratio 1.04, x and z periodic; the laminar similarity profile
    (similarity_profile) scaled to a 99 % thickness `delta`, plus `noise` x U
    seeded white noise on the velocity of the interior cells."""
    rng = np.random.default_rng(seed)
    s = stretching_for_ratio(ny, 1.04)
    nt, me, md, ms, y = wall_metrics(ny, width=Ly, stretching=s)
    W = 0.79 * AIR_W[0] + 0.21 * AIR_W[1]
    gamma = 1.4
    c_inf = math.sqrt(gamma * RGAS / W * T_inf)
    U = mach * c_inf
    Yb, ub, gb = similarity_profile(mach, T_wall / T_inf, gamma=gamma)
    Y99 = float(np.interp(0.99, ub, Yb))
    yb = Yb * (delta / Y99)                                          # physical wall distance
    uu = np.interp(y, yb, ub, right=1.0)
    T = T_inf * np.interp(y, yb, gb, right=1.0)
    nyg = ny + 2
    _, _, uRef = air_refs()
    vel = np.zeros((3, nz, nyg, nx))
    vel[0] = (U * uu)[None, :, None]
    vel[:, :, 1:-1, :] += noise * U * rng.standard_normal((3, nz, ny, nx))
    X = np.zeros((5, nz, nyg, nx))
    X[0], X[1] = 0.79, 0.21
    X = np.maximum(X, 1e-60)
    state = {"pressure": np.full((nz, nyg, nx), p / AIR_REFS["PRef"]),
             "temperature": np.broadcast_to((T / AIR_REFS["TRef"])[None, :, None], (nz, nyg, nx)).copy(),
             "MolarFracs": X, "velocity": vel / uRef}
    px = periodic_metrics(nx, Lx)
    pz = periodic_metrics(nz, Lz)
    set_metrics(state, [px, (nt, me, md, ms), pz], (nx, nyg, nz))
    meta = dict(kind="airmix", vorticityScale=U / uRef, stretching=s, U=U, **AIR_REFS)
    return state, meta


# ---- serving cases ----------------------------------------------------------------
def shock_tube(n, pressure_ratio, species=1):
    """1-D shock tube along x: n cells on [0, 1], diaphragm at 0.5, gas at rest,
    left pressure = pressure_ratio x right, equal temperatures; staggered
    Dirichlet ghosts at both ends (they keep the initial states).  species 1:
    calorically perfect gas (gamma 1.4, R 1, p_R = 1, T = 1); 5: air (p_R = 1e4
    Pa, T = 3000 K, X(N2, O2) = (0.79, 0.21)).  Returns (state, meta) with
    arrays shaped (…, 1, 1, n + 2)."""
    nx = n + 2
    xc = np.concatenate([[0.0], (np.arange(n) + 0.5) / n, [1.0]])
    left = xc < 0.5
    if species == 1:
        pR, T = 1.0, 1.0
        X = np.ones((1, 1, 1, nx))
        meta = dict(kind="constprop", R=1.0, gamma=1.4, prandtl=0.71, visc=1.0e-4, vorticityScale=1.0)
    else:
        pR, T = 1.0e4 / AIR_REFS["PRef"], 3000.0 / AIR_REFS["TRef"]
        X = np.zeros((5, 1, 1, nx))
        X[0], X[1] = 0.79, 0.21
        X = np.maximum(X, 1e-60)
        meta = dict(kind="airmix", vorticityScale=1.0, **AIR_REFS)
    p = np.where(left, pressure_ratio * pR, pR)[None, None, :]
    state = {"pressure": p.copy(), "temperature": np.full((1, 1, nx), T), "MolarFracs": X,
             "velocity": np.zeros((3, 1, 1, nx))}
    nt, me, md, ms, _ = wall_metrics(n, width=1.0, stretching=1e-3)   # uniform to 1e-7
    one = (np.zeros(1, np.int32), np.ones(1), np.ones(1), np.ones(1))
    set_metrics(state, [(nt, me, md, ms), one, one], (nx, 1, 1))
    return state, meta


def isentropic_vortex(n, width=10.0, beta=5.0, gamma=1.4):
    """2-D isentropic vortex advected by u = 1 on a periodic n x n square of
    side `width` (calorically perfect gas, R = 1; the reference's
    VortexAdvection2D test family).  Arrays shaped (…, 1, n, n)."""
    x = cell_centres(n, width) - 0.5 * width
    X, Y = np.meshgrid(x, x, indexing="xy")               # (y, x)
    r2 = X ** 2 + Y ** 2
    f = np.exp(0.5 * (1.0 - r2))
    u = 1.0 - beta / (2 * math.pi) * Y * f
    v = beta / (2 * math.pi) * X * f
    T = 1.0 - (gamma - 1.0) * beta ** 2 / (8 * gamma * math.pi ** 2) * f ** 2
    rho = T ** (1.0 / (gamma - 1.0))
    state = {"pressure": (rho * T)[None], "temperature": T[None], "MolarFracs": np.ones((1, 1, n, n)),
             "velocity": np.stack([u, v, np.zeros_like(u)])[:, None]}
    px = periodic_metrics(n, width)
    one = (np.zeros(1, np.int32), np.ones(1), np.ones(1), np.ones(1))
    set_metrics(state, [px, px, one], (n, n, 1))
    meta = dict(kind="constprop", R=1.0, gamma=gamma, prandtl=0.71, visc=1.0e-4, vorticityScale=1.0)
    return state, meta


# ---- wall-bounded direction -----------------------------------------------------
STD, L_S, LP1_S, LP2_S, RM3_S, RM2_S, RM1_S, R_S, L_C, LP1_C, RM2_C, RM1_C, R_C = range(13)


def tanh_minus_coords(n, width, stretching):
    """Cell centres of a TanhMinus-stretched direction with one staggered
    ghost at each end placed on the wall faces (src/prometeo_grid.rg:153-160,
    296-304)."""
    xi = (np.arange(n) + 0.5) / n
    x = width * (1.0 - np.tanh(stretching * (1.0 - xi)) / math.tanh(stretching))
    return np.concatenate([[0.0], x, [width]])


def wall_metrics(n_interior, width=1.0, stretching=1.0, collocated=False):
    """Node types and metrics of a direction with walls on both ends,
    staggered (L_S .. R_S, src/prometeo_metric_coeffs.h:55-67) or collocated
    (L_C, Lp1_C .. Rm2_C, Rm1_C, R_C): returns (nType, m_e, m_d, m_s,
    centres) of length n_interior + 2."""
    x = tanh_minus_coords(n_interior, width, stretching)
    n = len(x)
    nt = np.full(n, STD, np.int32)
    if collocated:
        nt[0], nt[1] = L_C, LP1_C
        nt[-1], nt[-2], nt[-3] = R_C, RM1_C, RM2_C
    else:
        nt[0], nt[1], nt[2] = L_S, LP1_S, LP2_S
        nt[-1], nt[-2], nt[-3], nt[-4] = R_S, RM1_S, RM2_S, RM3_S
    # face coordinates: midpoint between neighbouring centres (ghosts sit on faces)
    xf = np.empty(n + 1)
    xf[1:-1] = 0.5 * (x[:-1] + x[1:])
    xf[1] = x[0]
    xf[-2] = x[-1]
    xf[0] = x[0]
    xf[-1] = x[-1]
    me = np.empty(n)
    for i in range(n):
        lo, hi = xf[i], xf[i + 1]
        me[i] = 1.0 / (hi - lo) if hi > lo else 1.0 / (x[min(i + 1, n - 1)] - x[max(i - 1, 0)])
    md = np.empty(n)
    ms = np.empty(n)
    grad = {STD: (0.5, 0.5), L_S: (0.0, 2.0), LP1_S: (1.0, 0.5), RM1_S: (0.5, 1.0), R_S: (2.0, 0.0),
            L_C: (0.0, 1.0), R_C: (1.0, 0.0)}
    for i in range(n):
        g0, g1 = grad.get(int(nt[i]), (0.5, 0.5))
        xm, xp = x[max(i - 1, 0)], x[min(i + 1, n - 1)]
        md[i] = 1.0 / (g0 * (x[i] - xm) + g1 * (xp - x[i]))
        ms[i] = 1.0 / (x[i + 1] - x[i]) if i + 1 < n else ms[i - 1]
    return nt, me, md, ms, x
