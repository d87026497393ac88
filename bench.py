"""Benchmark of the HTR right-hand-side hot path on one MI355X.

Workload (one "step" = one full RHS of an RK stage): 256^3 periodic box of
5-species air (Dirichlet(8,2,.5,.5,.5) mass fractions per cell, T0 ~ U(2500,
7000) K with 5 % noise, p = 1e4 Pa, k^-5/3 solenoidal velocity at Mach_rms
0.4), hybrid TENO/KG Euler fluxes + multicomponent diffusion + chemistry
source, FP64, in the reference accumulation order.  Inputs are resident in
HBM before the timed region.

Prints one JSON line: metric/value (RHS cell-updates/s), the roofline of the
dominant kernel (HIP events on the library's stream) and a CPU baseline (the
oracle restatement on the host cores, bounded sample).
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

# Algorithmic bytes per cell (FP64 fields read once, Conserved_t written once):
#   full RHS (AirMix): 51 doubles read + 3 int32 node types + 3 bool sensors
#   + 9 doubles written = 495 B/cell
#   one Euler direction (AirMix): Conserved 9, Conserved_old 9, rho, SoS,
#   MassFracs 5, velocity 3, pressure, m_e (29 doubles), nType int32, sensor
#   bool, Conserved_t read-modify-write 9 doubles x 2 = 389 B/cell
#   one diffusion direction (AirMix): Conserved[0:5], T, X 5, rho, mu, lam,
#   Di 5, velocity 3, m_s, m_d, m_d1, m_d2 (25 doubles), 3 int32 node types,
#   Conserved_t RMW 144 B = 356 B/cell
#   chemistry source: rho, p, T, Y 5 (64 B) + Conserved_t[0:5] RMW (80 B) = 144 B/cell
#   zero Conserved_t: 72 B/cell
BYTES_PER_CELL = {"rhs": 495.0, "euler": 389.0, "diffusion": 356.0, "chem_source": 144.0, "zero": 72.0}


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
    return ap.parse_args()


def setup_gpu(n, seed):
    import numpy as np
    import torch
    import htr_amd as H
    from htr_amd import cases
    st, meta = cases.air_box(n, seed)
    mix = H.Mixture.airmix(meta["PRef"], meta["TRef"], meta["LRef"], meta["xref"])
    fs = H.FieldSet((n, n, n), 5)
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


def run_rhs(H, fs, mix, scheme, dt, events=None):
    """One full RHS, launched task by task so each launch can be timed."""
    import torch
    lib = H.lib()
    import ctypes as C
    g = C.byref(fs.grid())
    f = C.byref(fs.c())
    s = H.stream_ptr()
    marks = []

    def mark(name):
        if events is not None:
            e = torch.cuda.Event(enable_timing=True)
            e.record()
            marks.append((name, e))

    mark("start")
    fs["Conserved_t"].zero_()
    mark("zero")
    H.check(lib.htr_chem_source(g, f, mix.h, s))
    mark("chem_source")
    for d in (2, 1, 0):
        H.check(lib.htr_euler_flux(d, scheme, g, f, 1.0, 0.0, dt, mix.h, None, s))
        mark("euler%d" % d)
    for d in (2, 1, 0):
        H.check(lib.htr_diffusion_flux(d, g, f, mix.h, s))
        mark("diffusion%d" % d)
    if events is not None:
        events.append(marks)


def cpu_baseline(n, seed, scheme, budget_s):
    import numpy as np
    from htr_amd import cases
    from oracle import oracle as O
    st, meta = cases.air_box(n, seed)
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


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (see DESIGN.md, sharding)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import htr_amd as H
    H.init(0)
    scheme = {"hybrid": H.HTR_EULER_HYBRID, "TENO": H.HTR_EULER_TENO, "KG": H.HTR_EULER_KG}[a.scheme]
    n = a.n
    cells = n ** 3
    t_setup = time.perf_counter()
    fs, mix, dt, meta = setup_gpu(n, a.seed)
    print("setup %.1fs, n=%d, dt=%.3e, T0=%.0f K" % (time.perf_counter() - t_setup, n, dt, meta["T0"]),
          file=sys.stderr, flush=True)
    for _ in range(a.warmup):
        run_rhs(H, fs, mix, scheme, dt)
    torch.cuda.synchronize()
    events = []
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record()
    for _ in range(a.steps):
        run_rhs(H, fs, mix, scheme, dt, events)
    t1.record()
    torch.cuda.synchronize()
    total_ms = t0.elapsed_time(t1)
    ms_step = total_ms / a.steps
    # per-launch averages
    per = {}
    for marks in events:
        for (n0, e0), (n1, e1) in zip(marks[:-1], marks[1:]):
            per.setdefault(n1, []).append(e0.elapsed_time(e1))
    avg = {k: sum(v) / len(v) for k, v in per.items()}
    dom = max(avg, key=avg.get)
    kind = "euler" if dom.startswith("euler") else "diffusion" if dom.startswith("diffusion") else dom
    bpc = BYTES_PER_CELL[kind]
    achieved = bpc * cells / (avg[dom] * 1e-3) / 1e9
    kg_frac = float(torch.stack([fs["shockSensor%d" % d].double().mean() for d in range(3)]).mean())
    result = {
        "metric": "RHS cell-updates/s (256^3 5-species air, full FP64 RHS)",
        "value": cells / (ms_step * 1e-3),
        "unit": "cell-updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "air5_box_%d^3_rhs_%s" % (n, a.scheme), "grid": [n, n, n], "species": 5,
                   "scheme": a.scheme, "terms": "chem source + Euler ZYX + diffusion ZYX", "seed": a.seed},
        "kernel_ms": {k: round(v, 4) for k, v in sorted(avg.items())},
        "rhs_hbm_frac_495B": 495.0 * cells / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "smooth_sensor_frac": kg_frac,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "bytes_per_cell": bpc},
    }
    if not a.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        rate, reps, el = cpu_baseline(a.cpu_sample, a.seed, scheme, a.cpu_seconds)
        result["cpu_baseline"] = {"value": rate, "unit": "cell-updates/s", "cores": threads, "kind": "port",
                                  "sample": "%d^3 box of the same workload, %d full RHS in %.1f s"
                                            % (a.cpu_sample, reps, el)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
