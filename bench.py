"""Benchmark of the HTR right-hand-side hot path on one MI355X.

Workload (one "step" = one full RHS of an RK stage): 256^3 periodic box of
5-species air (Dirichlet(8,2,.5,.5,.5) mass fractions drawn per seed, T0 ~
U(2500, 7000) K with 5 % Gaussian noise, p = 1e4 Pa, k^-5/3 solenoidal
velocity at Mach_rms 0.4), hybrid TENO/KG Euler fluxes + multicomponent
diffusion + chemistry source, FP64, in the reference accumulation order.
Inputs are resident in HBM before the timed region.  --composition per-cell
draws the mass fractions independently in every cell (a TENO-heavy stress
reading).

Prints one JSON line: metric/value (RHS cell-updates/s), per-kernel times
from HIP events recorded by the library around each launch on its stream,
the roofline of the dominant kernel, the whole-RHS fraction of the HBM
roofline at the 495 B/cell contract, and a CPU baseline (the oracle
restatement on the host cores, bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "htr-solver_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6

# Compulsory (algorithmic) HBM bytes of one full RHS cell-update (AirMix):
# 51 doubles read once + 3 int32 node types + 3 bool sensors + Conserved_t
# (9 doubles) written once = 495 B.
RHS_BYTES_PER_CELL = 495.0

# Per-kernel algorithmic bytes per unit (AirMix, NS = 5, nEq = 9); the unit
# count of one launch comes from the workload (cells, faces, TENO faces).
#   euler_classify  per face: nType + own-cell sensor (5 B) read
#   euler_kg_species per cell (KG walk, species group): rho, u_n, Y 5 (56 B)
#                   + nType + sensor (5 B) + m_e (8 B) read, Conserved_t[0:5]
#                   read-modify-write (80 B) = 149 B
#   euler_kg_flow   per cell (KG walk, momentum/energy group): rho, u 3, H,
#                   p (48 B) + 13 B as above, Conserved_t[5:9] RMW (64 B) = 125 B
#   euler_roe       per TENO face: own-cell Roe inputs (88 B) + SoS (8 B)
#                   read, face record (22 doubles, 176 B) written = 272 B
#   euler_teno      per TENO face: face record (21 doubles, 168 B) read, the
#                   limiter test's rho Y_i, old rho Y_i and m_e of the face's two
#                   cells (22 doubles, 176 B) read, face flux (72 B) written =
#                   416 B; the stencil points are shared between faces and
#                   not counted; FP64-bound, see KERNEL_FLOPS
#   euler_divergence per cell: face flux 72 B + m_e 8 B read, Conserved_t
#                   read-modify-write 144 B = 224 B
#   diffusion_species per cell: rho Y 5, T, X 5, rho, lambda, D 5, m_s, m_d
#                   (20 doubles) + nType read, heat flux per face written,
#                   Conserved_t[0:5] read-modify-write (80 B) = 252 B
#   diffusion_stress per cell: mu, u 3, m_s, m_d and the transverse m_d 2
#                   (8 doubles) + 3 node types + heat flux read,
#                   Conserved_t[5:9] RMW (64 B) = 148 B
#   chem_source     per cell: rho, p, T, Y 5 (64 B) read, Conserved_t (9
#                   doubles, 72 B) written = 136 B (the source also clears
#                   the stage's Conserved_t)
#   enthalpy        per cell: rhoE, p, rho read, H written = 32 B
KERNEL_BYTES = {"euler_classify": 5.0, "euler_kg_species": 149.0, "euler_kg_flow": 125.0, "euler_roe": 272.0,
                "euler_teno": 416.0, "euler_divergence": 224.0, "diffusion_species": 252.0,
                "diffusion_stress": 148.0, "chem_source_kernel": 136.0, "enthalpy": 32.0}
# FP64 operations per unit of the compute-bound kernels (hand count of the
# source, FMA-free build: per component the species sums and split fluxes of
# six points ~600, two TENO reconstructions ~300; nine components plus the
# back-projection and limiter test ~150)
KERNEL_FLOPS = {"euler_teno": 8250.0}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--scheme", default="hybrid", choices=["hybrid", "TENO", "KG"])
    ap.add_argument("--cpu-sample", type=int, default=48, help="edge of the CPU-baseline box")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the per-cell-composition and boundary-layer readings of the default run")
    ap.add_argument("--composition", default="per-seed", choices=["per-seed", "per-cell"])
    ap.add_argument("--layout", default="soa", choices=["soa", "aos", "padded"],
                    help="vector fields structure-of-arrays, per point as the reference's VecNEq/VecNSp/Vec3, "
                         "or structure-of-arrays with padded (non power-of-two) strides")
    ap.add_argument("--workload", default="air", choices=["air", "bl"],
                    help="air: 256^3 periodic air box; bl: 512x256x128 Mach 6 boundary layer")
    ap.add_argument("--chem-n", type=int, default=32,
                    help="edge of the batched-chemistry sample (0: skip)")
    ap.add_argument("--chem-dt", type=float, default=1e-5)
    ap.add_argument("--chem-only", action="store_true",
                    help="only the batched-chemistry measurement (its own JSON line; --chem-n, --chem-dt)")
    ap.add_argument("--no-chem-tune", dest="chem_tune", action="store_false",
                    help="skip the launch-shape autotuner of the chemistry (default: each path's shape is chosen "
                         "by timing the candidates on a quarter of the cells)")
    ap.add_argument("--serve-requests", type=int, default=24, help="requests of the serving run (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=40.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--overlap", default="off", choices=["on", "off"],
                    help="hybrid full RHS with each direction's TENO passes on a side stream (on) or one stream")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy a 1 GiB FP64 tensor 5 times (reference kernel for HBM counters)")
    return ap.parse_args()


def setup_bl(seed):
    """Boundary-layer workload: 512 x 256 x 128 interior cells (y between two
    staggered wall ghosts), Mach 6 air."""
    import numpy as np
    import torch
    import htr_amd as H
    from htr_amd import cases
    st, meta = cases.boundary_layer(512, 256, 128, seed)
    mix = H.Mixture.airmix(meta["PRef"], meta["TRef"], meta["LRef"], meta["xref"])
    n = (512, 258, 128)
    fs = H.FieldSet(n, 5)
    for k, v in st.items():
        fs.t[k].copy_(torch.from_numpy(np.ascontiguousarray(v)).to(fs.t[k].dtype))
    del st
    H.update_properties_from_primitive(fs, mix)
    H.update_conserved_from_primitive(fs, mix)
    H.rk_init(fs, 9)
    H.update_ducros_sensor(fs, meta["vorticityScale"])
    for d in range(3):
        H.update_shock_sensor(fs, d, mix)
    grid = fs.grid((0, 1, 0), (511, 256, 127))
    r = H.calculate_max_spectral_radius(fs, mix, grid=grid)
    torch.cuda.synchronize()
    meta["T0"] = 220.0
    return fs, mix, 0.5 / r, meta, grid


def setup_gpu(n, seed, composition="per-seed", layout="soa"):
    import numpy as np
    import torch
    import htr_amd as H
    from htr_amd import cases
    st, meta = cases.air_box(n, seed, composition=composition)
    mix = H.Mixture.airmix(meta["PRef"], meta["TRef"], meta["LRef"], meta["xref"])
    fs = H.FieldSet((n, n, n), 5, layout=layout)
    for k, v in st.items():
        fs.t[k].copy_(torch.from_numpy(np.ascontiguousarray(v)).to(fs.t[k].dtype))
    del st
    H.update_properties_from_primitive(fs, mix)
    H.update_conserved_from_primitive(fs, mix)
    H.rk_init(fs, 9)
    H.update_ducros_sensor(fs, meta["vorticityScale"])
    for d in range(3):
        H.update_shock_sensor(fs, d, mix)
    r = H.calculate_max_spectral_radius(fs, mix)
    dt = 0.5 / r
    torch.cuda.synchronize()
    return fs, mix, dt, meta


def run_rhs(H, fs, mix, scheme, dt, grid=None):
    """One full RHS through the library's one-call entry point (htr_full_rhs:
    chemistry source clearing Conserved_t, total enthalpy once, Euler Z Y X,
    diffusion Z Y X, the reference accumulation order) on torch's current
    stream."""
    H.full_rhs(fs, scheme, (1.0, 0.0), dt, True, mix, grid=grid)


def face_counts(H, fs, mix, scheme, dt, n, grid=None):
    """TENO and limiter face counts per direction (one instrumented pass,
    outside the timed region)."""
    import torch
    import numpy as np
    fs2 = H.FieldSet(fs.n, 5, decisions=True)
    for k, v in fs.t.items():
        if k in fs2.t:
            fs2.t[k].copy_(v)
    out = {}
    for d in range(3):
        H.update_using_euler_flux(fs2, d, scheme, (1.0, 0.0), dt, mix, grid=grid, diag=True)
        dec = fs2["decisions%d" % d].cpu().numpy().astype(np.uint32)
        out[d] = (int(np.sum((dec & 1) == 0)), int(np.sum((dec & 2) != 0)))
    del fs2
    torch.cuda.empty_cache()
    return out


def cpu_baseline(n, seed, scheme, budget_s, composition="per-seed"):
    import numpy as np
    from htr_amd import cases
    from oracle import oracle as O
    st, meta = cases.air_box(n, seed, composition=composition)
    om = O.Mix.airmix(meta["PRef"], meta["TRef"], meta["LRef"], meta["xref"])
    g = O.Grid(n, n, n)
    f = O.Fields(g, 5)
    for k, v in st.items():
        f[k] = v
    O.update_properties(g, f, om)
    O.update_conserved(g, f, om)
    f["Conserved_old"] = f["Conserved"]
    O.ducros_sensor(g, f, 1.0)
    for d in range(3):
        O.shock_sensor(g, f, d, om)
    dt = 0.5 / O.max_spectral_radius(g, f, om)
    reps, t0 = 0, time.perf_counter()
    while True:
        O.full_rhs(g, f, scheme, 1.0, 0.0, dt, 1, om)
        reps += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    return reps * n ** 3 / el, reps, el


TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "rhs_air5_256_hbm_traffic.md")


def measured_traffic(launch):
    """HBM bytes per launch of `launch` on the default workload, from the
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (copy-calibrated), or None."""
    try:
        with open(TRAFFIC_TABLE) as fh:
            for line in fh:
                cols = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cols) == 6 and cols[0] == launch and cols[5].isdigit():
                    return int(cols[5])
    except OSError:
        pass
    return None


def kernel_family(name):
    for k in KERNEL_BYTES:
        if name.startswith(k):
            return k
    return None


def units_of(fam, name, nfaces_dir, cells, counts):
    """Units one launch of kernel `name` processes."""
    d = {"x": 0, "y": 1, "z": 2}.get(name[-1], 0)
    teno = counts[d][0] if counts else 0
    if fam == "euler_classify":
        return nfaces_dir[d]
    if fam in ("euler_roe", "euler_teno"):
        return teno
    return cells


def chemistry(n, dt, seed, cpu, threads, tune=True):
    """Batched implicit chemistry, 30 species / 180 reversible reactions
    (34 x 34 Jacobians): solves/s of four paths on n^3 cells -- FP64 with the
    reference's central-difference Jacobian (the plain FP64 path), FP32 LU +
    FP64 refinement with that Jacobian, and both precisions with the analytic
    Jacobian -- launch shapes chosen by the autotuner; accuracy of every path
    against the converged solution (the oracle at step tolerance 1e-13) on a
    4^3 sub-sample, per species; CPU baseline: the oracle on that sample."""
    import numpy as np
    import torch
    import htr_amd as H
    from htr_amd import cases
    st, meta = cases.chem30_box(n, seed)
    mix = H.Mixture.from_mechanism(meta["mech"])
    fs = H.FieldSet((n, n, n), 30, nsteps=True)
    for k, v in st.items():
        fs.t[k].copy_(torch.from_numpy(np.ascontiguousarray(v)).to(fs.t[k].dtype))
    H.update_properties_from_primitive(fs, mix)
    H.update_conserved_from_primitive(fs, mix)
    fs["Conserved_t"].zero_()
    init = {k: fs[k].clone() for k in ("Conserved", "temperature")}
    sub = fs.grid((0, 0, 0), (n - 1, n - 1, max(n // 4, 1) - 1))
    out = {"workload": "chem30_%d^3" % n, "cells": n ** 3, "dt": dt, "jacobian": "34x34 (nEq)",
           "mechanism": "CH4_30SpMix species, 180 synthetic reversible reactions (seed 0)"}
    AJ = H.HTR_CHEM_JAC_ANALYTIC
    paths = ((H.HTR_CHEM_FP64, "fp64"), (H.HTR_CHEM_FP32_IR, "fp32_ir"),
             (H.HTR_CHEM_FP64 | AJ, "fp64_aj"), (H.HTR_CHEM_FP32_IR | AJ, "fp32_ir_aj"))
    res = {}
    for prec, name in paths:
        if tune:
            best, table = H.chem_autotune(fs, mix, dt, prec, grid=sub,
                                          configs=[(16, 0), (32, 0), (64, 0), (64, 4)])
            print("chemistry %s: autotune %s -> %s" % (name, {k: round(v, 2) for k, v in table.items()}, best),
                  file=sys.stderr, flush=True)
        else:
            best, table = (64, 0), {}
            H.chem_launch_config(*best)
        for k in init:
            fs[k].copy_(init[k])
        fs["Conserved_t"].zero_()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        H.update_chemistry(fs, dt, mix, precision=prec, diag=True)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        steps = fs["nsteps"].cpu().numpy().ravel()
        res[name] = fs["Conserved"][:30].cpu().numpy().copy()
        print("chemistry %s: %d cells in %.1f ms" % (name, n ** 3, ms), file=sys.stderr, flush=True)
        out[name] = {"solves_per_s": n ** 3 / (ms * 1e-3), "ms": ms, "launch": {"lanes_per_cell": best[0], "blocks_per_cu": best[1]},
                     "autotune_ms": {"%dx%d" % k: round(v, 3) for k, v in table.items()},
                     "steps_mean": float(steps.mean()), "steps_p50": float(np.percentile(steps, 50)),
                     "steps_p99": float(np.percentile(steps, 99)), "steps_max": int(steps.max())}
        for k in init:
            fs[k].copy_(init[k])
    H.chem_launch_config(0, 0)
    for name in ("fp32_ir", "fp32_ir_aj", "fp64_aj"):
        d = np.abs(res[name] - res["fp64"])
        out[name + "_vs_fp64"] = {"max_abs": float(d.max()),
                                  "max_rel_per_species": [float(x) for x in d.reshape(30, -1).max(axis=1) /
                                                          np.abs(res["fp64"]).reshape(30, -1).max(axis=1)]}
    out["speedup_fp32_ir"] = out["fp32_ir_aj"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    out["speedup_fp32_ir_same_jacobian"] = {"fd": out["fp32_ir"]["solves_per_s"] / out["fp64"]["solves_per_s"],
                                            "analytic": out["fp32_ir_aj"]["solves_per_s"] / out["fp64_aj"]["solves_per_s"]}
    if cpu:
        # the converged solution of a 4^3 sub-sample (oracle, analytic
        # Jacobian, tolerance 1e-13) and the CPU baseline (oracle, the
        # reference algorithm) on the same cells
        from oracle import oracle as O
        m = O.Mix.from_mechanism(meta["mech"])
        ns = 4
        g = O.Grid(ns, ns, ns)
        sl = (slice(0, ns),) * 3

        def sample():
            f = O.Fields(g, 30)
            for k in ("Conserved", "temperature"):
                a = init[k].cpu().numpy()
                f[k] = a[(slice(None),) + sl] if a.ndim == 4 else a[sl]
            f["Conserved_t"] = 0.0
            return f
        f = sample()
        t0 = time.perf_counter()
        O.chem_update(g, f, dt, m)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": ns ** 3 / el, "unit": "solves/s", "cores": threads, "kind": "port",
                               "sample": "4^3 cells of the same workload, %.2f s" % el}
        fc = sample()
        O.chem_update(g, fc, dt, m, jacobian=O.JAC_ANALYTIC, tol=1e-13)
        conv = fc["Conserved"][:30]
        scale = np.abs(conv).reshape(30, -1).max(axis=1)
        acc = {}
        for name in res:
            d = np.abs(res[name][(slice(None),) + sl] - conv).reshape(30, -1).max(axis=1)
            acc[name] = {"max_abs": float(d.max()), "max_rel_per_species": [float(x) for x in d / scale]}
        out["error_vs_converged_4^3"] = acc
        print("chemistry error vs converged: %s" % {k: "%.2e" % v["max_abs"] for k, v in acc.items()},
              file=sys.stderr, flush=True)
    return out


def serving(nreq, rate, seed):
    """Continuous-batching request loop: a seeded Poisson stream of shock tubes
    (N ~ U{256..16384}, pressure ratio U[2, 20], 1 or 5 species) and 2-D
    periodic vortices (64^2 .. 1024^2), 20-60 SSP-RK3 steps each, admitted into
    packs by free capacity and HBM budget; p50/p99 of the per-step latency."""
    from htr_amd import serving as S
    S.warm_up()                                   # start-up, untimed: first launches load kernel code
    reqs = S.poisson_requests(nreq, rate, seed=seed, steps=(20, 60), vortex_sizes=(64, 128, 256, 512, 1024))
    srv = S.Server(budget_bytes=32 << 30, tube_capacity=1 << 16, vortex_columns=1024)
    stats = srv.run(reqs, progress=lambda m: print(m, file=sys.stderr, flush=True))
    stats.update({"arrival_rate_per_s": rate, "seed": seed,
                  "mix": {k: sum(1 for r in reqs if r.kind == k) for k in S.KINDS},
                  "unit": "cell-steps/s (one SSP-RK3 step of one cell)"})
    waits = [r.admitted - r.arrival for r in reqs]
    stats["admission_wait_ms_p50"] = float(sorted(waits)[len(waits) // 2] * 1e3)
    print("serving: %s" % {k: stats[k] for k in ("rounds", "wall_s", "step_latency_ms_p50", "step_latency_ms_p99")},
          file=sys.stderr, flush=True)
    return stats


def calibrate_copy():
    """A plain device-to-device copy of known size: FETCH_SIZE / WRITE_SIZE of
    this kernel, divided by its 1 GiB per launch, scale the counters of the hot
    kernels to bytes."""
    import torch
    n = 1 << 27
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 5,
                      "ms_per_copy": ms, "GBps_read_plus_write": 2 * n * 8 / (ms * 1e-3) / 1e9}), flush=True)


def measure_rhs(H, a, workload, composition, traffic_table=False):
    """Sets up the workload, counts its TENO faces, times a.steps full RHS
    after a.warmup; returns the reading (throughput, per-kernel times, the
    roofline of the dominant kernel) and frees the fields."""
    import torch
    scheme = {"hybrid": H.HTR_EULER_HYBRID, "TENO": H.HTR_EULER_TENO, "KG": H.HTR_EULER_KG}[a.scheme]
    n = a.n
    t_setup = time.perf_counter()
    if workload == "bl":
        fs, mix, dt, meta, grid = setup_bl(a.seed)
        cells = 512 * 256 * 128
        nfaces_dir = [(513 * 256 * 128), (512 * 257 * 128), (512 * 256 * 129)]
    else:
        fs, mix, dt, meta = setup_gpu(n, a.seed, composition, a.layout)
        grid = None
        cells = n ** 3
        nfaces_dir = [(n + 1) * n * n] * 3
    print("setup %.1fs, %s %s, dt=%.3e" % (time.perf_counter() - t_setup, workload, composition, dt),
          file=sys.stderr, flush=True)
    counts = face_counts(H, fs, mix, scheme, dt, n, grid)
    print("TENO faces per direction %s" % counts, file=sys.stderr, flush=True)
    H.rhs_overlap(a.overlap == "on")
    H.workspace_reserve(fs, mix, grid=grid)
    print("timing %d RHS steps" % (a.warmup + a.steps), file=sys.stderr, flush=True)
    for _ in range(a.warmup):
        run_rhs(H, fs, mix, scheme, dt, grid)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record()
    for _ in range(a.steps):
        run_rhs(H, fs, mix, scheme, dt, grid)
    t1.record()
    torch.cuda.synchronize()
    total_ms = t0.elapsed_time(t1)
    ms_step = total_ms / a.steps
    # per-kernel times (HIP events around every launch) from a second pass
    # with every launch on one stream: with the overlap, kernels share the
    # GPU and an event pair would time the sharing, not the kernel
    H.rhs_overlap(False)
    torch.cuda.synchronize()
    H.profile(True)
    H.profile_report()
    p0, p1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    p0.record()
    for _ in range(a.steps):
        run_rhs(H, fs, mix, scheme, dt, grid)
    p1.record()
    torch.cuda.synchronize()
    rep = H.profile_report()
    H.profile(False)
    ms_one_stream = p0.elapsed_time(p1) / a.steps
    per_launch = {k: ms / cnt for k, (cnt, ms) in rep.items()}
    per_step = {k: ms / a.steps for k, (cnt, ms) in rep.items()}
    dom = max(per_step, key=per_step.get)
    fam = kernel_family(dom)
    units = units_of(fam, dom, nfaces_dir, cells, counts) if fam else cells
    launch_ms = per_launch[dom]
    bpu = KERNEL_BYTES.get(fam, 0.0)
    gbs = bpu * units / (launch_ms * 1e-3) / 1e9
    if fam in KERNEL_FLOPS:
        # FP64-bound: the FP64 rate against the 78.6 TFLOP/s FP64 peak (the
        # MI355X FP64 vector and matrix peaks are the same figure)
        tf = KERNEL_FLOPS[fam] * units / (launch_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "kernel": dom, "achieved": tf, "peak": FP64_PEAK_TFS, "unit": "TFLOP/s",
                "frac": tf / FP64_PEAK_TFS, "traffic": None, "flops_per_unit": KERNEL_FLOPS[fam],
                "units_per_launch": units, "launch_ms": launch_ms,
                "note": "FP64 VALU-bound (no MFMA); peak = MI355X FP64 78.6 TFLOP/s",
                "hbm_achieved_GBps": gbs, "bytes_per_unit": bpu}
    else:
        roof = {"bound": "hbm", "kernel": dom, "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": gbs / HBM_PEAK_GBS, "traffic": None, "bytes_per_unit": bpu, "units_per_launch": units,
                "launch_ms": launch_ms}
    traffic = measured_traffic(dom) if traffic_table else None
    if traffic is not None:
        roof["traffic"] = traffic
        roof["traffic_source"] = os.path.relpath(TRAFFIC_TABLE, ROOT)
        roof["algorithmic_bytes"] = bpu * units
    faces = sum(nfaces_dir)
    out = {"value": cells / (ms_step * 1e-3), "ms_per_step": ms_step, "cells": cells,
           "rhs_overlap": a.overlap == "on", "ms_per_step_one_stream_profiled": ms_one_stream,
           "kernel_ms_per_step": {k: round(v, 4) for k, v in sorted(per_step.items())},
           "teno_face_frac": sum(c[0] for c in counts.values()) / faces,
           "limiter_faces": {d: c[1] for d, c in counts.items()},
           "rhs_roofline": {"bytes_per_cell": RHS_BYTES_PER_CELL,
                            "achieved": RHS_BYTES_PER_CELL * cells / (ms_step * 1e-3) / 1e9,
                            "frac": RHS_BYTES_PER_CELL * cells / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
           "roofline": roof}
    print("RHS %s %s: %.3f ms/step" % (workload, composition, ms_step), file=sys.stderr, flush=True)
    del fs
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return out


def main():
    a = parse()
    if a.calibrate_copy:
        calibrate_copy()
        return
    if a.chem_only:
        import htr_amd as H
        H.init(0)
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out = chemistry(max(a.chem_n, 4), a.chem_dt, a.seed, not a.no_cpu, threads, a.chem_tune)
        out.update({"metric": "implicit-chemistry solves/s (30 species, 34x34 Jacobians), FP32+IR vs FP64",
                    "value": out["fp32_ir_aj"]["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
                    "higher_is_better": True, "dtype": "f64 (FP32 LU + FP64 refinement)", "data": "synthetic"})
        print(json.dumps(out), flush=True)
        return
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one GPU's share of the "
              "domain is the unit of work; see the sharding section of DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import htr_amd as H
    H.init(0)
    n = a.n
    default_workload = (a.workload == "air" and n == 256 and a.composition == "per-seed" and a.scheme == "hybrid"
                        and a.layout == "soa")
    main_r = measure_rhs(H, a, a.workload, a.composition, traffic_table=default_workload)
    result = {
        "metric": "RHS cell-updates/s (%s, 5-species air, full FP64 RHS)" % ("256^3 box" if a.workload == "air" else "512x256x128 boundary layer"),
        "value": main_r["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": main_r["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": ("air5_box_%d^3_rhs_%s_%s" % (n, a.scheme, a.composition)) if a.workload == "air"
                   else "air5_bl_512x256x128_rhs_%s" % a.scheme,
                   "grid": [n, n, n] if a.workload == "air" else [512, 256, 128],
                   "species": 5, "scheme": a.scheme, "composition": a.composition if a.workload == "air" else None,
                   "terms": "chem source + Euler ZYX + diffusion ZYX", "seed": a.seed,
                   "layout": a.layout if a.workload == "air" else "soa"},
    }
    for k in ("kernel_ms_per_step", "teno_face_frac", "limiter_faces", "rhs_roofline", "roofline", "rhs_overlap",
              "ms_per_step_one_stream_profiled"):
        result[k] = main_r[k]
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if default_workload and not a.no_extras:
        # the same box with the mass fractions drawn per cell (TENO-heavy)
        pc = measure_rhs(H, a, "air", "per-cell")
        result["per_cell_composition"] = {k: pc[k] for k in ("value", "ms_per_step", "teno_face_frac",
                                                             "rhs_roofline", "roofline", "kernel_ms_per_step")}
        # the boundary-layer workload (config 4)
        bl = measure_rhs(H, a, "bl", None)
        result["boundary_layer"] = {k: bl[k] for k in ("value", "ms_per_step", "teno_face_frac", "rhs_roofline",
                                                       "roofline")}
    if a.chem_n > 0:
        result["chemistry"] = chemistry(a.chem_n, a.chem_dt, a.seed, not a.no_cpu, threads, a.chem_tune)
    if a.serve_requests > 0:
        result["serving"] = serving(a.serve_requests, a.serve_rate, a.seed)
    if not a.no_cpu:
        scheme = {"hybrid": H.HTR_EULER_HYBRID, "TENO": H.HTR_EULER_TENO, "KG": H.HTR_EULER_KG}[a.scheme]
        rate, reps, el = cpu_baseline(a.cpu_sample, a.seed, scheme, a.cpu_seconds, a.composition)
        result["cpu_baseline"] = {"value": rate, "unit": "cell-updates/s", "cores": threads, "kind": "port",
                                  "sample": "%d^3 box of the same workload, %d full RHS in %.1f s"
                                            % (a.cpu_sample, reps, el)}
        # the same restatement on one core (a smaller box, same workload)
        from oracle import oracle as O
        O.set_threads(1)
        rate1, reps1, el1 = cpu_baseline(24, a.seed, scheme, a.cpu_seconds / 2, a.composition)
        O.set_threads(threads)
        result["cpu_baseline"]["one_core"] = {"value": rate1, "cores": 1,
                                              "sample": "24^3 box, %d full RHS in %.1f s" % (reps1, el1)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
