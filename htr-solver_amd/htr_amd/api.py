"""Task-level API mirroring the reference's Legion tasks of the hot path.

Each function takes a FieldSet (device tensors of the Fluid_columns members,
src/prometeo_types.h:48-114), a Mixture and the task's scalar arguments, and
runs the corresponding C-ABI entry point on the current HIP stream.
"""
import ctypes as C

import numpy as np
import torch

from . import Diag, Fields, Grid, View, check, lib, HTR_CHEM_FP64
from . import HTR_EULER_HYBRID, HTR_EULER_TENO, HTR_EULER_TENOA, HTR_EULER_TENOLAD, HTR_EULER_KG

AIR_SPECIES = ("N2", "O2", "NO", "N", "O")

__all__ = ["Mixture", "FieldSet", "make_grid", "AIR_SPECIES",
           "update_properties_from_primitive", "update_conserved_from_primitive",
           "update_primitive_from_conserved", "update_ducros_sensor", "update_shock_sensor",
           "update_using_euler_flux", "update_using_hybrid_euler_flux", "update_using_teno_euler_flux",
           "update_using_skew_symmetric_euler_flux", "update_using_diffusion_flux",
           "add_chemistry_sources", "update_chemistry", "calculate_max_spectral_radius",
           "max_spectral_radius_noalloc",
           "rk_update", "rk_init", "full_rhs", "stream_ptr", "profile", "profile_report",
           "chem_launch_config", "chem_autotune", "rk_stage", "full_rhs_cfl", "workspace_reserve", "set_bc",
           "BC_SIDES", "update_using_flux_nscbc_inflow", "update_using_flux_nscbc_outflow",
           "update_using_flux_nscbc_far_field",
           "calculate_max_mach_number", "rhs_overlap"]


def stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class Mixture:
    """A device mixture image (LoadMixtureTask): ConstPropMix, AirMix or a
    generic multicomponent mechanism."""

    def __init__(self, handle, nspec):
        self.h = handle
        self.nSpec = nspec

    @property
    def nEq(self):
        return self.nSpec + 4

    @classmethod
    def constprop(cls, R, gamma, prandtl, visc_model=0, p0=0.0, p1=0.0, p2=0.0):
        h = C.c_void_p()
        check(lib().htr_mix_create_constprop(C.byref(h), R, gamma, prandtl, visc_model, p0, p1, p2))
        return cls(h, 1)

    @classmethod
    def airmix(cls, PRef, TRef, LRef, xref):
        h = C.c_void_p()
        idx = (C.c_int * len(xref))(*[AIR_SPECIES.index(k) for k in xref])
        val = (C.c_double * len(xref))(*[xref[k] for k in xref])
        check(lib().htr_mix_create_airmix(C.byref(h), PRef, TRef, LRef, len(xref), idx, val))
        return cls(h, 5)

    @classmethod
    def from_mechanism(cls, mech):
        ns = len(mech["W"])
        h = C.c_void_p()
        a = lambda k, t: np.ascontiguousarray(mech[k], dtype=t)
        W, cpT, nPoly = a("W", np.float64), a("cpT", np.float32), a("nPoly", np.int32)
        cpH, cpM, cpL, geom, dc = a("cpH", np.float64), a("cpM", np.float64), a("cpL", np.float64), \
            a("geom", np.int32), a("dc", np.float64)
        dp, fp, ip = C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_int)
        check(lib().htr_mix_create_species(C.byref(h), ns, W.ctypes.data_as(dp), cpT.ctypes.data_as(fp),
                                           nPoly.ctypes.data_as(ip), cpH.ctypes.data_as(dp),
                                           cpM.ctypes.data_as(dp), cpL.ctypes.data_as(dp),
                                           geom.ctypes.data_as(ip), dc.ctypes.data_as(dp)))
        m = cls(h, ns)
        for r in mech["reactions"]:
            ei = np.array([e[0] for e in r["educts"]], np.int32)
            en = np.array([e[1] for e in r["educts"]], np.float32)
            pi = np.array([e[0] for e in r["products"]], np.int32)
            pn = np.array([e[1] for e in r["products"]], np.float32)
            ti = np.array([e[0] for e in r.get("thirdb", [])], np.int32)
            te = np.array([e[1] for e in r.get("thirdb", [])], np.float32)
            check(lib().htr_mix_add_reaction(h, r["kind"], r["A"], r["n"], r["EovR"], int(r["has_backward"]),
                                             len(ei), ei.ctypes.data_as(ip), en.ctypes.data_as(fp),
                                             len(pi), pi.ctypes.data_as(ip), pn.ctypes.data_as(fp),
                                             len(ti), ti.ctypes.data_as(ip) if len(ti) else None,
                                             te.ctypes.data_as(fp) if len(te) else None))
        xr = np.ascontiguousarray(mech["XiRef"], np.float64)
        check(lib().htr_mix_finalize(h, mech["PRef"], mech["TRef"], mech["LRef"], xr.ctypes.data_as(dp)))
        return m

    def refs(self):
        out = (C.c_double * 9)()
        check(lib().htr_mix_refs(self.h, out))
        keys = ("rhoRef", "ieRef", "iuRef", "iCpRef", "imuRef", "ilamRef", "iDiRef", "iwiRef", "MixWRef")
        return dict(zip(keys, list(out)))

    def __del__(self):
        try:
            if self.h:
                lib().htr_mix_destroy(self.h)
        except Exception:
            pass


def make_grid(n, lo=(0, 0, 0), hi=None, bounds_lo=(0, 0, 0)):
    """Grid over a box of n = (nx, ny, nz) cells whose first cell has absolute
    coordinates bounds_lo; exec rect [lo, hi] in box-relative indices."""
    g = Grid()
    hi = hi if hi is not None else tuple(x - 1 for x in n)
    for d in range(3):
        g.bounds.lo[d] = bounds_lo[d]
        g.bounds.hi[d] = bounds_lo[d] + n[d] - 1
        g.exec.lo[d] = bounds_lo[d] + lo[d]
        g.exec.hi[d] = bounds_lo[d] + hi[d]
    return g


SCALARS = ("rho", "mu", "lam", "SoS", "pressure", "temperature", "DucrosSensor")
VECTORS = {"Di": "ns", "MassFracs": "ns", "MolarFracs": "ns", "velocity": 3,
           "Conserved": "neq", "Conserved_old": "neq", "Conserved_t": "neq", "Conserved_t_old": "neq"}
PROFILES = {"MolarFracs_profile": "ns", "velocity_profile": 3, "temperature_profile": 0,
            "dudtBoundary": 3, "dTdtBoundary": 0}


class FieldSet:
    """Device fields of one tile.  Every vector field is presented as a
    (ncomp, nz, ny, nx) tensor, x fastest.  layout "soa": components far apart
    (one block per component); "aos": the reference's per-point arrays
    (VecNEq, VecNSp, Vec3 of src/prometeo_types.h), stored (nz, ny, nx, ncomp)
    and presented through a permuted view."""

    def __init__(self, n, nspec, device="cuda", bounds_lo=(0, 0, 0), decisions=False, nsteps=False,
                 layout="soa", flags=False, profiles=False):
        nx, ny, nz = n
        self.n = tuple(n)
        self.ns = nspec
        self.bounds_lo = tuple(bounds_lo)
        kw = dict(device=device)
        self.t = {}
        if layout == "padded":
            # structure of arrays with padded strides (one element-stride
            # triple for every field): rows of nx + 16 elements (128-byte
            # aligned in FP64), planes of ny + 1 rows, components one line
            # apart, and each field's base staggered by a few cache lines, so
            # that no stride is a power of two and the same cell of different
            # rows, planes and fields falls in different cache sets
            sy = nx + 16
            sz = sy * (ny + 1)
            sc = sz * nz + 16
            nfield = [0]

            def alloc(nc, dtype, fill):
                stag = (nfield[0] % 8) * 40 * 8 // torch.empty((), dtype=dtype).element_size()
                nfield[0] += 1
                base = torch.full((max(nc, 1) * sc + stag,), fill, dtype=dtype, **kw)
                if nc == 0:
                    return base.as_strided((nz, ny, nx), (sz, sy, 1), stag)
                return base.as_strided((nc, nz, ny, nx), (sc, sz, sy, 1), stag)
        else:
            def alloc(nc, dtype, fill):
                shape = (nz, ny, nx) if nc == 0 else (nc, nz, ny, nx)
                if nc and layout == "aos":
                    return torch.full((nz, ny, nx, nc), fill, dtype=dtype, **kw).permute(3, 0, 1, 2)
                return torch.full(shape, fill, dtype=dtype, **kw)
        for k in SCALARS:
            self.t[k] = alloc(0, torch.float64, 0)
        for k, c in VECTORS.items():
            nc = {"ns": nspec, "neq": nspec + 4}.get(c, c)
            self.t[k] = alloc(nc, torch.float64, 0)
        for d in range(3):
            self.t["nType%d" % d] = alloc(0, torch.int32, 0)
            for k in ("m_e", "m_d", "m_s"):
                self.t["%s%d" % (k, d)] = alloc(0, torch.float64, 1)
            self.t["shockSensor%d" % d] = alloc(0, torch.uint8, 1)
            if decisions:
                self.t["decisions%d" % d] = alloc(0, torch.int32, 0)
        if nsteps:
            self.t["nsteps"] = alloc(0, torch.int32, 0)
        if flags:
            self.t["flags"] = alloc(0, torch.int32, 0)
        if profiles:
            for k, c in PROFILES.items():
                nc = {"ns": nspec}.get(c, c)
                self.t[k] = alloc(nc, torch.float64, 0)
        self._c = None

    def __getitem__(self, k):
        return self.t[k]

    def __setitem__(self, k, v):
        if isinstance(v, np.ndarray):
            v = torch.from_numpy(np.ascontiguousarray(v))
        self.t[k].copy_(v if torch.is_tensor(v) else torch.tensor(v, dtype=self.t[k].dtype))

    def view(self, name):
        t = self.t[name]
        v = View()
        if t.dim() == 4:
            sc, sz, sy, sx = t.stride()
        else:
            sz, sy, sx = t.stride()
            sc = 0
        lo = self.bounds_lo
        off = lo[0] * sx + lo[1] * sy + lo[2] * sz
        v.ptr = t.data_ptr() - off * t.element_size()
        v.sx, v.sy, v.sz, v.sc = sx, sy, sz, sc
        return v

    def c(self):
        f = Fields()
        for d in range(3):
            f.nType[d] = self.view("nType%d" % d)
            f.m_e[d] = self.view("m_e%d" % d)
            f.m_d[d] = self.view("m_d%d" % d)
            f.m_s[d] = self.view("m_s%d" % d)
            f.shockSensor[d] = self.view("shockSensor%d" % d)
        for k in SCALARS + tuple(VECTORS):
            setattr(f, k, self.view(k))
        for k in PROFILES:
            if k in self.t:
                setattr(f, k, self.view(k))
        return f

    def diag(self):
        dg = Diag()
        for d in range(3):
            if "decisions%d" % d in self.t:
                dg.decisions[d] = self.view("decisions%d" % d)
        if "nsteps" in self.t:
            dg.nsteps = self.view("nsteps")
        if "flags" in self.t:
            dg.flags = self.view("flags")
        return dg

    def grid(self, lo=(0, 0, 0), hi=None):
        return make_grid(self.n, lo, hi, self.bounds_lo)


def _g(fs, grid):
    return C.byref(grid if grid is not None else fs.grid())


# ---- tasks --------------------------------------------------------------------
def update_properties_from_primitive(fs, mix, grid=None, stream=None):
    check(lib().htr_update_properties(_g(fs, grid), C.byref(fs.c()), mix.h, stream_ptr(stream)))


def update_conserved_from_primitive(fs, mix, grid=None, stream=None):
    check(lib().htr_update_conserved(_g(fs, grid), C.byref(fs.c()), mix.h, stream_ptr(stream)))


def update_primitive_from_conserved(fs, mix, grid=None, stream=None):
    check(lib().htr_update_primitive(_g(fs, grid), C.byref(fs.c()), mix.h, stream_ptr(stream)))


def update_ducros_sensor(fs, vorticity_scale, grid=None, stream=None):
    check(lib().htr_ducros_sensor(_g(fs, grid), C.byref(fs.c()), vorticity_scale, stream_ptr(stream)))


def update_shock_sensor(fs, direction, mix, grid=None, stream=None):
    check(lib().htr_shock_sensor(direction, _g(fs, grid), C.byref(fs.c()), mix.h, stream_ptr(stream)))


def update_using_euler_flux(fs, direction, scheme, RK_coeffs, deltaTime, mix, grid=None, stream=None, diag=False):
    dg = C.byref(fs.diag()) if diag else None
    check(lib().htr_euler_flux(direction, scheme, _g(fs, grid), C.byref(fs.c()), RK_coeffs[0], RK_coeffs[1],
                               deltaTime, mix.h, dg, stream_ptr(stream)))


def update_using_hybrid_euler_flux(fs, direction, RK_coeffs, deltaTime, mix, **kw):
    update_using_euler_flux(fs, direction, HTR_EULER_HYBRID, RK_coeffs, deltaTime, mix, **kw)


def update_using_teno_euler_flux(fs, direction, op, RK_coeffs, deltaTime, mix, **kw):
    scheme = {"TENO": HTR_EULER_TENO, "TENOA": HTR_EULER_TENOA, "TENOLAD": HTR_EULER_TENOLAD}[op]
    update_using_euler_flux(fs, direction, scheme, RK_coeffs, deltaTime, mix, **kw)


def update_using_skew_symmetric_euler_flux(fs, direction, mix, **kw):
    update_using_euler_flux(fs, direction, HTR_EULER_KG, (0.0, 0.0), 0.0, mix, **kw)


def update_using_diffusion_flux(fs, direction, mix, grid=None, stream=None):
    check(lib().htr_diffusion_flux(direction, _g(fs, grid), C.byref(fs.c()), mix.h, stream_ptr(stream)))


def add_chemistry_sources(fs, mix, grid=None, stream=None):
    check(lib().htr_chem_source(_g(fs, grid), C.byref(fs.c()), mix.h, stream_ptr(stream)))


def update_chemistry(fs, deltaTime, mix, precision=HTR_CHEM_FP64, grid=None, stream=None, diag=False):
    dg = C.byref(fs.diag()) if diag else None
    check(lib().htr_chem_update(_g(fs, grid), C.byref(fs.c()), deltaTime, precision, mix.h, dg, stream_ptr(stream)))


def calculate_max_spectral_radius(fs, mix, grid=None, stream=None):
    out = C.c_double(0.0)
    check(lib().htr_max_spectral_radius(_g(fs, grid), C.byref(fs.c()), mix.h, C.byref(out), stream_ptr(stream)))
    return out.value


def rk_update(fs, nEq, stage, dt, grid=None, stream=None):
    check(lib().htr_rk_update(_g(fs, grid), C.byref(fs.c()), nEq, stage, dt, stream_ptr(stream)))


def rk_init(fs, nEq, grid=None, stream=None):
    check(lib().htr_rk_init(_g(fs, grid), C.byref(fs.c()), nEq, stream_ptr(stream)))


def rhs_overlap(on):
    """Scheduling of the hybrid full RHS: True runs each direction's TENO
    passes on a library side stream, overlapping the chemistry source and the
    earlier directions' KG walks (same result to the bit); False (the default)
    runs every launch on the caller's stream."""
    check(lib().htr_rhs_overlap(1 if on else 0))


def full_rhs(fs, scheme, RK_coeffs, deltaTime, chem_source, mix, grid=None, stream=None):
    check(lib().htr_full_rhs(_g(fs, grid), C.byref(fs.c()), scheme, RK_coeffs[0], RK_coeffs[1], deltaTime,
                             int(chem_source), mix.h, stream_ptr(stream)))


def max_spectral_radius_noalloc(fs, mix, out, grid=None, stream=None):
    """calculate_max_spectral_radius into the one-element float64 device
    tensor out, without allocation or host read-back."""
    check(lib().htr_max_spectral_radius_noalloc(_g(fs, grid), C.byref(fs.c()), mix.h, C.c_void_p(out.data_ptr()),
                                                stream_ptr(stream)))


def full_rhs_cfl(fs, scheme, RK_coeffs, cfl, scale, radius, chem_source, mix, grid=None, stream=None):
    """full_rhs with the stage's deltaTime = (cfl / radius) * scale formed on
    the device; radius is a one-element float64 device tensor."""
    check(lib().htr_full_rhs_cfl(_g(fs, grid), C.byref(fs.c()), scheme, RK_coeffs[0], RK_coeffs[1], cfl, scale,
                                 C.c_void_p(radius.data_ptr()), int(chem_source), mix.h, stream_ptr(stream)))


def rk_stage(fs, stage, mix, deltaTime=0.0, cfl=0.0, radius=None, grid=None, stream=None, diag=False):
    """Fused SSP-RK3 stage: update, primitive recompute and properties in one
    pass per cell; with radius (device tensor) the step is cfl / radius."""
    dg = C.byref(fs.diag()) if diag else None
    rp = C.c_void_p(radius.data_ptr()) if radius is not None else None
    check(lib().htr_rk_stage(_g(fs, grid), C.byref(fs.c()), stage, deltaTime, cfl, rp, mix.h, dg,
                             stream_ptr(stream)))


BC_SIDES = ("xNeg", "xPos", "yNeg", "yPos", "zNeg", "zPos")


def set_bc(fs, kind, side, grid, mix, Pbc=0.0, stream=None):
    """Ghost-cell primitives of one boundary (SetDirichletBC /
    SetAdiabaticWallBC / SetIsothermalWallBC); side is an index or a name of
    BC_SIDES; grid's execution rect is the boundary's ghost cells."""
    sd = BC_SIDES.index(side) if isinstance(side, str) else int(side)
    check(lib().htr_set_bc(int(kind), sd, C.byref(grid), C.byref(fs.c()), float(Pbc), mix.h, stream_ptr(stream)))


def _side(side):
    return BC_SIDES.index(side) if isinstance(side, str) else int(side)


def update_using_flux_nscbc_inflow(fs, side, mix, grid, stream=None):
    """UpdateUsingFluxNSCBCInflow{Minus,Plus}Side: LODI corrections of
    Conserved_t in the boundary cells of one side (grid's execution rect),
    reading the inflow profile's dudtBoundary and dTdtBoundary."""
    check(lib().htr_nscbc_inflow(_side(side), C.byref(grid), C.byref(fs.c()), mix.h, stream_ptr(stream)))


def update_using_flux_nscbc_outflow(fs, side, mix, MaxMach, LengthScale, PInf, grid, stream=None):
    """UpdateUsingFluxNSCBCOutflow{Minus,Plus}Side: LODI corrections of
    Conserved_t in the boundary cells of one side, the incoming wave relaxed
    towards PInf over LengthScale; MaxMach from calculate_max_mach_number."""
    check(lib().htr_nscbc_outflow(_side(side), C.byref(grid), C.byref(fs.c()), float(MaxMach), float(LengthScale),
                                  float(PInf), mix.h, stream_ptr(stream)))


def update_using_flux_nscbc_far_field(fs, side, mix, PInf, MaxMach, LengthScale, grid, stream=None):
    """UpdateUsingFluxNSCBCFarField{Minus,Plus}Side: LODI corrections of
    Conserved_t towards the far-field state of the profile fields."""
    check(lib().htr_nscbc_farfield(_side(side), C.byref(grid), C.byref(fs.c()), float(PInf), float(MaxMach),
                                   float(LengthScale), mix.h, stream_ptr(stream)))


def calculate_max_mach_number(fs, direction, grid=None, stream=None):
    """CalculateMaxMachNumber: max |u_dir|/SoS over the rect (host value)."""
    out = C.c_double(0.0)
    check(lib().htr_max_mach(int(direction), _g(fs, grid), C.byref(fs.c()), C.byref(out), stream_ptr(stream)))
    return out.value


def workspace_reserve(fs, mix, grid=None, stream=None):
    """Sizes the stream's library workspace for this tile (no allocation in
    later calls on it)."""
    check(lib().htr_workspace_reserve(_g(fs, grid), C.byref(fs.c()), mix.h, stream_ptr(stream)))


def profile(enable=True):
    """Per-launch HIP-event timing inside the library (htr_profile)."""
    check(lib().htr_profile(1 if enable else 0))


def profile_report():
    """{kernel name: (launches, total ms)} of the launches since the last
    report (waits for them)."""
    buf = C.create_string_buffer(1 << 16)
    check(lib().htr_profile_report(buf, len(buf)))
    out = {}
    for line in buf.value.decode().splitlines():
        name, n, ms = line.rsplit(" ", 2)
        out[name] = (int(n), float(ms))
    return out


def chem_launch_config(lanes_per_cell, blocks_per_cu=0):
    """Launch shape of the implicit chemistry: lanes of a wave per cell (8, 16,
    32, 64; 0 = default) and resident one-wave blocks per CU (0 = as many as
    the LDS holds)."""
    check(lib().htr_chem_launch_config(int(lanes_per_cell), int(blocks_per_cu)))


CHEM_CONFIGS = [(8, 0), (16, 0), (32, 0), (64, 0), (64, 2), (64, 4)]


def chem_autotune(fs, mix, deltaTime, precision=HTR_CHEM_FP64, grid=None, configs=None, repeats=1):
    """Times update_chemistry over `grid` for each launch shape (lanes per
    cell, resident blocks per CU) on a copy of the cell states (restored after every
    trial, so the caller's fields are unchanged), applies the fastest and
    returns (best, {shape: ms})."""
    names = ("Conserved", "temperature", "Conserved_t", "Conserved_t_old")
    saved = {k: fs[k].clone() for k in names}
    times = {}
    for cfg in (configs or CHEM_CONFIGS):
        chem_launch_config(*cfg)
        best = None
        for _ in range(repeats):
            for k in names:
                fs[k].copy_(saved[k])
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record()
            update_chemistry(fs, deltaTime, mix, precision=precision, grid=grid)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1)
            best = ms if best is None else min(best, ms)
        times[cfg] = best
    for k in names:
        fs[k].copy_(saved[k])
    top = min(times, key=times.get)
    chem_launch_config(*top)
    return top, times
