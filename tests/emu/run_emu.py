"""Child process of tests/test_emu_kernels.py: runs small hot-path cases
through the library's kernel source compiled for the CPU (the HIP emulation of
tests/emu/hip, AddressSanitizer + UBSan) and compares them with the oracle.

On the host both sides use glibc's libm and no FMA contraction, so every case
must agree bit for bit (the GPU parity tests allow 1e-12 for the device's own
transcendental functions).  Prints one line per case, "CASE <name> <max
abs difference> <cells>", and exits non-zero on the first mismatch.
"""
import ctypes
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "htr-solver_amd"), os.path.dirname(HERE)]

import numpy as np  # noqa: E402
import torch  # noqa: E402

import htr_amd as H  # noqa: E402
import htr_amd.api as A  # noqa: E402
from oracle import oracle as O  # noqa: E402
from hotpath import air_state, complete_state, gpu_mix, nscbc_state, oracle_fields, tgv_state, wall_state, \
    dt_cfl  # noqa: E402

# host tensors stand for device memory; the emulated library ignores streams
A.stream_ptr = lambda stream=None: ctypes.c_void_p(0)


def to_host_fs(state, nspec, n, **kw):
    fs = H.FieldSet(n, nspec, device="cpu", **kw)
    for k, v in state.items():
        if k in fs.t:
            fs.t[k].copy_(torch.from_numpy(np.ascontiguousarray(v)).to(fs.t[k].dtype))
    return fs


def report(name, got, ref, cells):
    d = float(np.max(np.abs(np.asarray(got, np.float64) - np.asarray(ref, np.float64))))
    print("CASE %s %.3e %d" % (name, d, cells), flush=True)
    if not d == 0.0:
        sys.exit(3)


def euler(st, meta, n, om, d, scheme, label, RK=(1.0, 0.0), dt=1e-3, lo=(0, 0, 0), hi=None):
    g, f = oracle_fields(st, om.nSpec, n)
    g = g.with_rect(lo, hi if hi else tuple(x - 1 for x in n))
    f["Conserved_t"] = 0.0
    dec_o = np.zeros(f["rho"].shape, np.uint32)
    O.euler_flux(g, f, d, scheme, RK[0], RK[1], dt, om, dec_o)
    fs = to_host_fs(st, om.nSpec, n, decisions=True)
    fs["Conserved_t"] = 0.0
    H.update_using_euler_flux(fs, d, scheme, RK, dt, gpu_mix(meta), grid=fs.grid(lo, hi), diag=True)
    report("euler_%s_d%d" % (label, d), fs["Conserved_t"].numpy(), f["Conserved_t"], int(np.prod(n)))
    dg = fs["decisions%d" % d].numpy().astype(np.uint32)
    mask = np.zeros(dg.shape, bool)
    lo3, hi3 = lo[::-1], (hi if hi else tuple(x - 1 for x in n))[::-1]
    mask[lo3[0]:hi3[0] + 1, lo3[1]:hi3[1] + 1, lo3[2]:hi3[2] + 1] = True
    flips = int(np.sum((dg != dec_o) & mask))
    print("CASE decisions_%s_d%d %d %d" % (label, d, flips, int(mask.sum())), flush=True)
    if flips:
        sys.exit(4)


def main():
    H.init(0)
    t0 = time.time()
    which = sys.argv[1:] or ["euler", "limiter", "diffusion", "rhs", "stage", "wall", "bc", "nscbc"]

    st, meta = air_state(8, seed=4)
    n = (8, 8, 8)
    st, om = complete_state(st, meta, n)
    if "euler" in which:
        for d in range(3):
            euler(st, meta, n, om, d, O.SCHEME_HYBRID, "hybrid")
        euler(st, meta, n, om, 1, O.SCHEME_TENOA, "tenoa")
        euler(st, meta, n, om, 2, O.SCHEME_TENOLAD, "tenolad")
        euler(st, meta, n, om, 0, O.SCHEME_KG, "kg")
        # a sub-rect that starts inside the box (segment and row offsets)
        euler(st, meta, n, om, 1, O.SCHEME_HYBRID, "subrect", lo=(1, 2, 3), hi=(6, 7, 5))
    if "pipeline" in which:
        # few TENO blocks (HTR_TENO_BLOCKS, set by the caller): every block
        # takes several chunks, so the loader's gather of chunk k + 1 overlaps
        # the components of chunk k
        st3, meta3 = air_state(12, seed=6)
        n3 = (12, 12, 12)
        st3, om3 = complete_state(st3, meta3, n3)
        for d in range(3):
            euler(st3, meta3, n3, om3, d, O.SCHEME_TENO, "pipeline")
    if "limiter" in which:
        st2 = dict(st)
        # a nearly vanishing species next to a rich one: the positivity
        # limiter fires on the TENO faces between them
        c = st2["MassFracs"].copy()
        c[2, :, :, ::2] = 1e-14
        c /= c.sum(axis=0, keepdims=True)
        st2["MassFracs"] = c
        X = st2["MolarFracs"].copy()
        X[2, :, :, ::2] = 1e-14
        X /= X.sum(axis=0, keepdims=True)
        st2["MolarFracs"] = X
        st2, om2 = complete_state(st2, meta, n)
        for d in range(3):
            euler(st2, meta, n, om2, d, O.SCHEME_TENO, "limiter", RK=(0.5, 0.5), dt=5e-2)
    if "diffusion" in which:
        for d in range(3):
            g, f = oracle_fields(st, 5, n)
            f["Conserved_t"] = 0.0
            O.diffusion_flux(g, f, d, om)
            fs = to_host_fs(st, 5, n)
            fs["Conserved_t"] = 0.0
            H.update_using_diffusion_flux(fs, d, gpu_mix(meta))
            report("diffusion_d%d" % d, fs["Conserved_t"].numpy(), f["Conserved_t"], 512)
    if "rhs" in which:
        # the full RHS (hybrid: with each direction's TENO passes on a side
        # stream of the library, and with every launch on one stream): whole
        # box, a sub-rect, pure KG, a box with staggered walls in y, the TGV
        dt = dt_cfl(st, meta, n)
        for label, scheme, lo, hi, ov in (("full_rhs", O.SCHEME_HYBRID, (0, 0, 0), (7, 7, 7), True),
                                          ("full_rhs_one_stream", O.SCHEME_HYBRID, (0, 0, 0), (7, 7, 7), False),
                                          ("full_rhs_subrect", O.SCHEME_HYBRID, (1, 2, 3), (6, 7, 5), True),
                                          ("full_rhs_kg", O.SCHEME_KG, (0, 0, 0), (7, 7, 7), True)):
            g, f = oracle_fields(st, 5, n)
            g = g.with_rect(lo, hi)
            O.full_rhs(g, f, scheme, 1.0, 0.0, dt, 1, om)
            fs = to_host_fs(st, 5, n)
            H.rhs_overlap(ov)
            H.full_rhs(fs, scheme, (1.0, 0.0), dt, True, gpu_mix(meta), grid=fs.grid(lo, hi))
            H.rhs_overlap(False)
            report(label, fs["Conserved_t"].numpy(), f["Conserved_t"], 512)
        stw, metaw, nw = wall_state(8, 6, 8)
        stw, omw = complete_state(stw, metaw, nw)
        dtw = dt_cfl(stw, metaw, nw) * 0.5
        g, f = oracle_fields(stw, 5, nw)
        g = g.with_rect((0, 1, 0), (7, 6, 7))
        O.full_rhs(g, f, O.SCHEME_HYBRID, 1.0, 0.0, dtw, 1, omw)
        fs = to_host_fs(stw, 5, nw)
        H.full_rhs(fs, 0, (1.0, 0.0), dtw, True, gpu_mix(metaw), grid=fs.grid((0, 1, 0), (7, 6, 7)))
        report("full_rhs_wall", fs["Conserved_t"].numpy(), f["Conserved_t"], 512)
        stt, metat = tgv_state(8)
        nt = (8, 8, 8)
        stt, omt = complete_state(stt, metat, nt)
        g, f = oracle_fields(stt, 1, nt)
        O.full_rhs(g, f, O.SCHEME_HYBRID, 1.0, 0.0, 1e-3, 0, omt)
        fs = to_host_fs(stt, 1, nt)
        H.full_rhs(fs, 0, (1.0, 0.0), 1e-3, False, gpu_mix(metat))
        report("full_rhs_tgv", fs["Conserved_t"].numpy(), f["Conserved_t"], 512)
    if "stage" in which:
        # per-seed composition: a state that stays physical over CFL-0.4
        # steps (the per-cell white-noise composition does not)
        st, meta = air_state(8, seed=0, composition="per-seed")
        st, om = complete_state(st, meta, n)
        gm = gpu_mix(meta)
        r = H.calculate_max_spectral_radius(to_host_fs(st, 5, n), gm)
        radius = torch.tensor([r], dtype=torch.float64)
        a = to_host_fs(st, 5, n, flags=True)
        b = to_host_fs(st, 5, n)
        for fs in (a, b):
            H.rk_init(fs, 9)
        for stage, c3 in enumerate((1.0, 0.25)):
            RK = [(1.0, 0.0), (0.75, 0.25)][stage]
            H.full_rhs_cfl(a, 0, RK, 0.4, c3, radius, False, gm)
            H.full_rhs(b, 0, RK, (0.4 / r) * c3, False, gm)
            H.rk_stage(a, stage, gm, cfl=0.4, radius=radius, diag=True)
            H.rk_update(b, 9, stage, 0.4 / r)
            H.update_primitive_from_conserved(b, gm)
            H.update_properties_from_primitive(b, gm)
            for k in ("Conserved", "temperature", "pressure", "rho", "mu", "lam", "Di", "SoS", "MassFracs"):
                report("stage%d_%s" % (stage, k), a[k].numpy(), b[k].numpy(), 512)
            fl = int(a["flags"].abs().max())
            print("CASE stage%d_flags %d 512" % (stage, fl), flush=True)
            if fl != 0:
                sys.exit(5)
    if "wall" in which:
        for colloc in (False, True):
            stw, metaw, nw = wall_state(8, 6, 8, collocated=colloc)
            stw, omw = complete_state(stw, metaw, nw)
            dtw = dt_cfl(stw, metaw, nw) * 0.5
            tag = "colloc" if colloc else "wall"
            for scheme, label in ((O.SCHEME_HYBRID, "hybrid"), (O.SCHEME_TENO, "teno"), (O.SCHEME_KG, "kg")):
                euler(stw, metaw, nw, omw, 1, scheme, tag + "_" + label, dt=dtw, lo=(0, 1, 0), hi=(7, 6, 7))
            for d in range(3):
                g, f = oracle_fields(stw, 5, nw)
                g = g.with_rect((0, 1, 0), (7, 6, 7))
                f["Conserved_t"] = 0.0
                O.diffusion_flux(g, f, d, omw)
                fs = to_host_fs(stw, 5, nw)
                fs["Conserved_t"] = 0.0
                H.update_using_diffusion_flux(fs, d, gpu_mix(metaw), grid=fs.grid((0, 1, 0), (7, 6, 7)))
                report("%s_diffusion_d%d" % (tag, d), fs["Conserved_t"].numpy(), f["Conserved_t"], 512)
    if "bc" in which:
        prof = {"MolarFracs_profile": st["MolarFracs"][:, ::-1].copy(),
                "velocity_profile": 0.1 * st["velocity"][:, ::-1].copy(),
                "temperature_profile": st["temperature"][::-1].copy() * 1.1}
        for side in range(6):
            d, neg = side // 2, side % 2 == 0
            lo, hi = [0, 0, 0], [7, 7, 7]
            lo[d] = hi[d] = 0 if neg else 7
            for kind in (0, 1, 2, 3, 4):
                g, f = oracle_fields(dict(st, **prof), 5, n)
                gr = g.with_rect(tuple(lo), tuple(hi))
                O.set_bc(gr, f, kind, side, 0.9, om)
                O.update_properties(gr, f, om)
                O.update_conserved(gr, f, om)
                fs = to_host_fs(dict(st, **prof), 5, n, profiles=True)
                grid = fs.grid(tuple(lo), tuple(hi))
                H.set_bc(fs, kind, side, grid, gpu_mix(meta), Pbc=0.9)
                H.update_properties_from_primitive(fs, gpu_mix(meta), grid=grid)
                H.update_conserved_from_primitive(fs, gpu_mix(meta), grid=grid)
                for k in ("temperature", "pressure", "rho", "Conserved"):
                    report("bc%d_kind%d_%s" % (side, kind, k), fs[k].numpy(), f[k], 64)
    if "nscbc" in which:
        sn, metan, nn = nscbc_state(6)
        sn, omn = complete_state(sn, metan, nn, sensors=False)
        gmn = gpu_mix(metan)
        for side in range(6):
            d, neg = side // 2, side % 2 == 0
            lo, hi = [0, 0, 0], [5, 5, 5]
            lo[d] = hi[d] = 0 if neg else 5
            g, f = oracle_fields(sn, 5, nn)
            fs = to_host_fs(sn, 5, nn, profiles=True)
            mm = O.max_mach(g, f, d)
            report("max_mach_d%d_side%d" % (d, side), H.calculate_max_mach_number(fs, d), mm, 216)
            for kind, Mmax in (("inflow", None), ("outflow", 0.5), ("outflow_sup", max(mm, 1.0)),
                               ("farfield", -0.5), ("farfield_sup", -max(mm, 1.0))):
                f["Conserved_t"] = 0.0
                fs["Conserved_t"] = 0.0
                gr, grid = g.with_rect(tuple(lo), tuple(hi)), fs.grid(tuple(lo), tuple(hi))
                if Mmax is None:
                    O.nscbc_inflow(gr, f, side, omn)
                    H.update_using_flux_nscbc_inflow(fs, side, gmn, grid)
                elif Mmax > 0:
                    O.nscbc_outflow(gr, f, side, Mmax, 0.8, 0.95, omn)
                    H.update_using_flux_nscbc_outflow(fs, side, gmn, Mmax, 0.8, 0.95, grid)
                else:
                    O.nscbc_farfield(gr, f, side, 0.95, -Mmax, 0.8, omn)
                    H.update_using_flux_nscbc_far_field(fs, side, gmn, 0.95, -Mmax, 0.8, grid)
                report("nscbc_%s_side%d" % (kind, side), fs["Conserved_t"].numpy(), f["Conserved_t"], 36)
    print("DONE %.1f s" % (time.time() - t0), flush=True)


if __name__ == "__main__":
    main()
