"""Shared builders for the parity tests: one state, two copies (oracle numpy
fields and device FieldSet), and the normalised comparison of the reference
parity contract (max |gpu - oracle| / max |oracle| per conserved component)."""
import math

import numpy as np

from oracle import oracle as O
from htr_amd import cases

TOL = 1e-12          # FP64 parity bound of the hot path (normalised, per component)


def oracle_mix(meta):
    if meta["kind"] == "mechanism":
        return O.Mix.from_mechanism(meta["mech"])
    if meta["kind"] == "constprop":
        return O.Mix.constprop_mix(meta["R"], meta["gamma"], meta["prandtl"], 0, meta["visc"])
    return O.Mix.airmix(meta["PRef"], meta["TRef"], meta["LRef"], meta["xref"])


def gpu_mix(meta):
    import htr_amd as H
    if meta["kind"] == "mechanism":
        return H.Mixture.from_mechanism(meta["mech"])
    if meta["kind"] == "constprop":
        return H.Mixture.constprop(meta["R"], meta["gamma"], meta["prandtl"], 0, meta["visc"])
    return H.Mixture.airmix(meta["PRef"], meta["TRef"], meta["LRef"], meta["xref"])


def oracle_fields(state, nspec, n):
    g = O.Grid(*n)
    f = O.Fields(g, nspec)
    for k, v in state.items():
        if k in f.a:
            f[k] = v
    return g, f


def complete_state(state, meta, n, sensors=True):
    """Properties, conserved variables and shock sensors by the oracle, so that
    the flux tests feed identical inputs to both sides."""
    m = oracle_mix(meta)
    g, f = oracle_fields(state, m.nSpec, n)
    O.update_properties(g, f, m)
    O.update_conserved(g, f, m)
    f["Conserved_old"] = f["Conserved"]
    if sensors:
        O.ducros_sensor(g, f, meta.get("vorticityScale", 1.0))
        for d in range(3):
            O.shock_sensor(g, f, d, m)
    out = dict(state)
    for k in ("rho", "mu", "lam", "Di", "SoS", "MassFracs", "Conserved", "Conserved_old", "DucrosSensor",
              "shockSensor0", "shockSensor1", "shockSensor2"):
        out[k] = f[k].copy()
    return out, m


def to_gpu(state, nspec, n, **kw):
    import torch
    from htr_amd import FieldSet
    fs = FieldSet(n, nspec, **kw)
    for k, v in state.items():
        if k in fs.t:
            fs.t[k].copy_(torch.from_numpy(np.ascontiguousarray(v)).to(fs.t[k].dtype))
    return fs


def flux_scale(state, d):
    """Per-component magnitude of one face flux term, |Q_k| m (|u_n| + c): the
    rounding floor of a flux difference.  Used as the normalisation of a
    component whose reference divergence vanishes identically (e.g. the z mass
    flux of the w = 0 Taylor-Green vortex), where max |ref| is round-off."""
    Q = np.asarray(state["Conserved"])
    wave = np.abs(state["velocity"][d]) + state["SoS"]
    m = state["m_e%d" % d]
    return [float(np.max(np.abs(Q[k]) * m * wave)) for k in range(Q.shape[0])]


FLOOR_FRACTION = 1e-3


def norm_err(gpu, ref, mask=None, floor=None):
    """Per-component max |gpu-ref| / max(max |ref|, FLOOR_FRACTION floor_k)
    (components on axis 0).  floor_k is the single-face flux magnitude from
    flux_scale: a divergence below 1e-3 of it is dominated by cancellation
    between two O(floor_k) face fluxes, whose rounding residue (~1e-16
    floor_k per face) is the only thing left to compare."""
    gpu = np.asarray(gpu, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    if ref.ndim == 3:
        gpu, ref = gpu[None], ref[None]
    errs = []
    for c in range(ref.shape[0]):
        scale = np.max(np.abs(ref[c]))
        if floor is not None:
            scale = max(scale, FLOOR_FRACTION * floor[c])
        d = np.abs(gpu[c] - ref[c])
        if mask is not None:
            d = np.where(mask, d, 0.0)
        errs.append(float(np.max(d) / scale) if scale > 0 else float(np.max(d)))
    return errs


def dt_cfl(state, meta, n, cfl=0.5):
    m = oracle_mix(meta)
    g, f = oracle_fields(state, m.nSpec, n)
    return cfl / O.max_spectral_radius(g, f, m)


def air_state(n, seed, composition="per-cell"):
    st, meta = cases.air_box(n, seed, composition=composition)
    return st, meta


def tgv_state(n):
    return cases.tgv(n)


def wall_state(nx, ny, nz, seed=3, collocated=False):
    """Air box with y bounded by walls on a stretched grid (staggered node
    types L_S .. R_S, or collocated L_C .. R_C); x and z periodic."""
    st, meta = cases.air_box(max(nx, nz), seed)
    rng = np.random.default_rng(seed)
    n = (nx, ny + 2, nz)
    out = {}
    for k, v in st.items():
        if k.startswith(("nType", "m_")):
            continue
        a = v[..., :nz, :, :nx] if v.ndim == 4 else v[:nz, :, :nx]
        reps = math.ceil((ny + 2) / a.shape[-2])
        a = np.concatenate([a] * reps, axis=-2)[..., : ny + 2, :]
        out[k] = np.ascontiguousarray(a)
    # boundary layer-like profile in y
    nt, me, md, ms, y = cases.wall_metrics(ny, width=1.0, stretching=1.5, collocated=collocated)
    prof = np.tanh(4 * y / y[-1])
    out["velocity"] = out["velocity"] * prof[None, None, :, None]
    px = cases.periodic_metrics(nx)
    pz = cases.periodic_metrics(nz)
    cases.set_metrics(out, [px, (nt, me, md, ms), pz], n)
    return out, meta, n


def nscbc_state(n, seed=5):
    """Air box bounded by collocated boundary nodes (L_C .. R_C) in all three
    directions on stretched grids, with a mean flow through it, a few
    supersonic cells on every face, random inflow-profile time derivatives
    (dudtBoundary, dTdtBoundary) and a far-field state in the profile fields:
    the inputs of the NSCBC tasks."""
    st, meta = cases.air_box(n, seed, composition="per-cell")
    rng = np.random.default_rng(seed)
    dims = [cases.wall_metrics(n - 2, width=1.0, stretching=1.2 + 0.1 * d, collocated=True)[:4] for d in range(3)]
    cases.set_metrics(st, dims, (n, n, n))
    st = {k: np.ascontiguousarray(v) for k, v in st.items()}
    u = st["velocity"]
    scale = np.sqrt(np.mean(u * u))
    u = 0.5 * u + np.array([0.9, -0.6, 0.4])[:, None, None, None] * scale
    # supersonic spots (mach ~3 along each direction)
    for d in range(3):
        idx = rng.integers(0, n, size=(6, 3))
        for z, y, x in idx:
            u[d, z, y, x] = (1 if rng.random() < 0.5 else -1) * 6.0 * scale
    st["velocity"] = u
    st["dudtBoundary"] = rng.normal(size=(3, n, n, n)) * scale
    st["dTdtBoundary"] = rng.normal(size=(n, n, n)) * 0.1 * np.mean(st["temperature"])
    # far-field state: a perturbed copy of the mean flow
    X = st["MolarFracs"]
    st["MolarFracs_profile"] = (X.mean(axis=(1, 2, 3))[:, None, None, None] * (1 + 0.1 * rng.random(X.shape)))
    st["MolarFracs_profile"] /= st["MolarFracs_profile"].sum(axis=0)
    st["velocity_profile"] = u * (1 + 0.05 * rng.normal(size=u.shape))
    st["temperature_profile"] = st["temperature"] * (1 + 0.05 * rng.random((n, n, n)))
    return st, meta, (n, n, n)
