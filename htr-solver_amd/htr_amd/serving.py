"""Continuous-batching request loop for small independent cases.

Requests (1-D shock tubes of 1 or 5 species, 2-D periodic isentropic vortices)
arrive over time and are advanced together, one SSP-RK3 step at a time, by
the library's kernels.  Cases of one class share a *pack*: one device field
set holding every resident case side by side along x, so that one launch per
kernel advances all of them:

  * a shock tube keeps its two staggered ghost cells; the node types of its
    wall-adjacent cells confine every stencil to the tube (as at a wall of
    the reference solver), so its neighbours in the pack never enter its
    update; the ghosts hold the Dirichlet states and are restored after every
    RK stage;
  * a periodic vortex gets HALO columns on each side in x, refreshed from its
    own periodic images before every stage; HALO = 6 covers the chain
    face flux <- hybrid sensor at p-1 <- Ducros sensor at p-3 <- velocity at
    p-4, the widest dependency of one cell's update (5 cells); along y the
    pack is periodic over exactly the vortex height.

A pack's x extent is a fixed capacity handed out in segments (first fit);
admission requires a free segment and, for a new pack, room in the HBM budget.
Requests that arrive while others are running are admitted at the next step
boundary, as segments free up.  Each pack runs on its own stream, so the packs'
launches overlap on the device; a pack's library workspace is sized when the
pack is made, so no step allocates.  A case's results do not depend on its
neighbours: the same case alone and in a mix gives the same fields, and they
equal a run of the case on its own box (the oracle test).

Per RK stage a pack issues: the Ducros and shock sensors, the chemistry
source (5 species), the Euler and diffusion sweeps of its directions, the
fused stage update (update + primitives + properties in one pass,
htr_rk_stage) and one index copy per field that refreshes every ghost and
halo cell of every resident case at once.
"""
import time

import numpy as np
import torch

from . import api as A
from . import cases

HALO = 6
KINDS = ("tube1", "tube5", "vortex")


class Request:
    def __init__(self, rid, kind, n, steps, arrival=0.0, pressure_ratio=10.0):
        self.id, self.kind, self.n, self.steps = rid, kind, int(n), int(steps)
        self.arrival, self.pressure_ratio = arrival, pressure_ratio
        self.done_steps = 0
        self.step_ms = []
        self.admitted = None
        self.admitted_round = None
        self.finished = None
        self.result = None

    def state(self):
        if self.kind == "vortex":
            return cases.isentropic_vortex(self.n)
        return cases.shock_tube(self.n, self.pressure_ratio, 1 if self.kind == "tube1" else 5)

    def width(self):
        """x extent the case occupies in a pack."""
        return self.n + 2 * HALO if self.kind == "vortex" else self.n + 2

    def cells(self):
        return self.n * self.n if self.kind == "vortex" else self.n


# fixed time step per class (normalised units): CFL 0.5 on the finest grid of
# the class at its strongest wave (tubes: N = 16384, pressure ratio 20)
DT = {"tube1": 0.5 / 16384 / 1.5, "tube5": 0.5 / 16384 / 6.0, "vortex": 0.5 * (10.0 / 1024) / 2.4}


def mixture(kind):
    st, meta = (cases.isentropic_vortex(8) if kind == "vortex"
                else cases.shock_tube(8, 2.0, 1 if kind == "tube1" else 5))
    if meta["kind"] == "constprop":
        return A.Mixture.constprop(meta["R"], meta["gamma"], meta["prandtl"], 0, meta["visc"])
    return A.Mixture.airmix(meta["PRef"], meta["TRef"], meta["LRef"], meta["xref"])


# per-cell state fields a ghost or halo cell carries (metrics and node types
# are static; Conserved_old and Conserved_t are stage temporaries)
STATE = ("Conserved", "temperature", "pressure", "MolarFracs", "velocity", "rho", "mu", "lam", "Di", "SoS",
         "MassFracs")


class Pack:
    """Resident cases of one class (kind, ny) side by side along x."""

    def __init__(self, kind, ny, capacity, device="cuda"):
        self.kind, self.ny, self.cap = kind, ny, capacity
        self.nspec = 5 if kind == "tube5" else 1
        self.fs = A.FieldSet((capacity, ny, 1), self.nspec, device=device)
        self.mix = mixture(kind)
        self.dt = DT[kind]
        self.dirs = (0, 1) if kind == "vortex" else (0,)
        self.free = [(0, capacity)]                 # (start, length), sorted
        self.resident = {}                          # id -> (request, x0)
        self.ghost = {}                             # id -> (flat indices, saved state) of a tube's two ghosts
        self.stream = torch.cuda.Stream() if device == "cuda" else None
        self._bnd = None                            # boundary copy plan (rebuilt on admission/release)
        # pad cells: a quiet state so that unused columns stay finite
        self._fill_pad(0, capacity)
        if self.stream is not None:
            # the library workspace of this stream, sized once for the pack
            A.workspace_reserve(self.fs, self.mix, stream=self.stream)
            torch.cuda.synchronize()

    def nbytes(self):
        return sum(t.numel() * t.element_size() for t in self.fs.t.values())

    def _fill_pad(self, x0, w):
        st, _ = (cases.isentropic_vortex(8) if self.kind == "vortex"
                 else cases.shock_tube(8, 1.0, self.nspec))
        sl = (Ellipsis, slice(x0, x0 + w))
        for k in ("pressure", "temperature"):
            self.fs[k][sl] = float(st[k].ravel()[0])
        self.fs["MolarFracs"][sl] = torch.as_tensor(st["MolarFracs"][..., :1, :1, :1].reshape(-1, 1, 1, 1))
        self.fs["velocity"][sl] = 0.0
        for d in range(3):
            self.fs["nType%d" % d][sl] = 0
        A.update_properties_from_primitive(self.fs, self.mix, grid=self._grid(x0, w))
        A.update_conserved_from_primitive(self.fs, self.mix, grid=self._grid(x0, w))

    def _grid(self, x0, w):
        return A.make_grid(self.fs.n, (x0, 0, 0), (x0 + w - 1, self.ny - 1, 0))

    def try_admit(self, req):
        w = req.width()
        for i, (s, l) in enumerate(self.free):
            if l >= w:
                self.free[i] = (s + w, l - w)
                if self.free[i][1] == 0:
                    self.free.pop(i)
                self._place(req, s)
                return True
        return False

    def _flat(self, x, y):
        return np.asarray(y) * self.cap + np.asarray(x)

    def _place(self, req, x0):
        st, _ = req.state()
        w = req.width()
        off = HALO if self.kind == "vortex" else 0
        sl = (Ellipsis, slice(x0 + off, x0 + off + (req.n if self.kind == "vortex" else w)))
        for k, v in st.items():
            if k in self.fs.t:
                self.fs[k][sl] = torch.as_tensor(np.ascontiguousarray(v)).to(self.fs[k].dtype)
        g = self._grid(x0, w)
        if self.kind == "vortex":
            # halos: periodic images of every field, metrics and node types included
            a, b = x0 + HALO, x0 + HALO + req.n
            for k in self.fs.t:
                t = self.fs[k]
                t[..., a - HALO:a] = t[..., b - HALO:b]
                t[..., b:b + HALO] = t[..., a:a + HALO]
        A.update_properties_from_primitive(self.fs, self.mix, grid=g)
        A.update_conserved_from_primitive(self.fs, self.mix, grid=g)
        if self.kind != "vortex":
            idx = torch.as_tensor(self._flat([x0, x0 + w - 1], [0, 0]), device=self.fs["rho"].device)
            self.ghost[req.id] = (idx, {k: self._cols(k)[:, idx].clone() for k in STATE})
        self.resident[req.id] = (req, x0)
        self._bnd = None

    def _cols(self, k):
        """Field k as (components, cells) over the pack's plane."""
        t = self.fs[k]
        return t.reshape(-1, self.cap * self.ny) if t.dim() == 4 else t.reshape(1, -1)

    def _plan(self):
        """Flat indices of every ghost and halo cell of the resident cases and
        of the cells their values come from (vortex halos: the periodic
        images; tube ghosts: the saved Dirichlet states)."""
        dev = self.fs["rho"].device
        if self.kind == "vortex":
            dst, src = [], []
            ys = np.arange(self.ny)
            for req, x0 in self.resident.values():
                a, b = x0 + HALO, x0 + HALO + req.n
                for h in range(HALO):
                    dst.append(self._flat(a - HALO + h, ys)); src.append(self._flat(b - HALO + h, ys))
                    dst.append(self._flat(b + h, ys)); src.append(self._flat(a + h, ys))
            dst = torch.as_tensor(np.concatenate(dst), device=dev)
            src = torch.as_tensor(np.concatenate(src), device=dev)
            return dst, src, None
        ids = list(self.ghost)
        dst = torch.cat([self.ghost[r][0] for r in ids])
        saved = {k: torch.cat([self.ghost[r][1][k] for r in ids], dim=1) for k in STATE}
        return dst, None, saved

    def _boundaries(self):
        """Dirichlet ghosts of the tubes, periodic halos of the vortices: one
        index copy per state field for the whole pack."""
        if not self.resident:
            return
        if self._bnd is None:
            self._bnd = self._plan()
        dst, src, saved = self._bnd
        for k in STATE:
            c = self._cols(k)
            c[:, dst] = c[:, src] if src is not None else saved[k]

    def release(self, rid):
        req, x0 = self.resident.pop(rid)
        self.ghost.pop(rid, None)
        self._bnd = None
        w = req.width()
        self._fill_pad(x0, w)
        self.free.append((x0, w))
        self.free.sort()
        merged = []
        for s, l in self.free:
            if merged and merged[-1][0] + merged[-1][1] == s:
                merged[-1] = (merged[-1][0], merged[-1][1] + l)
            else:
                merged.append((s, l))
        self.free = merged

    def extract(self, req, x0):
        """The case's conserved variables (interior cells) as a host array."""
        if self.kind == "vortex":
            return self.fs["Conserved"][..., x0 + HALO:x0 + HALO + req.n].cpu().numpy().copy()
        return self.fs["Conserved"][..., x0 + 1:x0 + 1 + req.n].cpu().numpy().copy()

    def step(self):
        """One SSP-RK3 step of every resident case (src/prometeo.rg:1407-1470):
        per stage the sensors, the RHS with the stage's deltaTime
        dt RK_C[stage][3], the fused update + primitives + properties, then
        the boundary values; one launch per kernel for the whole pack."""
        fs, mix, dt = self.fs, self.mix, self.dt
        neq = self.nspec + 4
        ctx = torch.cuda.stream(self.stream) if self.stream is not None else _null()
        with ctx:
            A.rk_init(fs, neq)
            for stage in range(3):
                A.update_ducros_sensor(fs, 1.0)
                for d in self.dirs:
                    A.update_shock_sensor(fs, d, mix)
                fs["Conserved_t"].zero_()
                if self.nspec > 1:
                    A.add_chemistry_sources(fs, mix)
                RK = ((1.0, 0.0), (0.75, 0.25), (1.0 / 3.0, 2.0 / 3.0))[stage]
                c3 = (1.0, 0.25, 2.0 / 3.0)[stage]
                for d in reversed(self.dirs):
                    A.update_using_euler_flux(fs, d, A.HTR_EULER_HYBRID, RK, dt * c3, mix)
                for d in reversed(self.dirs):
                    A.update_using_diffusion_flux(fs, d, mix)
                A.rk_stage(fs, stage, mix, deltaTime=dt)
                self._boundaries()


class _null:
    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False


class Server:
    """Admits requests by arrival time and HBM budget, advances all resident
    cases one step per round, retires finished ones."""

    def __init__(self, budget_bytes=16 << 30, tube_capacity=1 << 17, vortex_columns=4096, device="cuda",
                 max_packs_per_class=None):
        self.budget = budget_bytes
        self.max_packs = max_packs_per_class
        self.tube_cap = tube_capacity
        self.vortex_cols = vortex_columns
        self.device = device
        self.packs = {}

    def _pack_for(self, req):
        key = (req.kind, req.n if req.kind == "vortex" else 1)
        for p in self.packs.get(key, []):
            if p.try_admit(req):
                return p
        if self.max_packs is not None and len(self.packs.get(key, [])) >= self.max_packs:
            return None
        cap = self.vortex_cols if req.kind == "vortex" else self.tube_cap
        cap = max(cap, req.width())
        used = sum(p.nbytes() for ps in self.packs.values() for p in ps)
        est = cap * key[1] * (60 if req.kind == "tube5" else 40) * 8
        if used + est > self.budget:
            return None
        p = Pack(req.kind, key[1], cap, self.device)
        self.packs.setdefault(key, []).append(p)
        return p if p.try_admit(req) else None

    def run(self, requests, keep_results=False, clock=time.perf_counter, progress=None):
        """Serves the requests (sorted by arrival, seconds after start); returns
        statistics.  Per-step latency of a case = wall time of the round that
        advanced it."""
        pending = sorted(requests, key=lambda r: r.arrival)
        queue = []
        t0 = clock()
        cell_steps = 0
        rounds = 0
        while pending or queue or any(p.resident for ps in self.packs.values() for p in ps):
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.pop(0))
            still = []
            for req in queue:
                p = self._pack_for(req)
                if p is None:
                    still.append(req)
                else:
                    req.admitted = clock() - t0
                    req.admitted_round = rounds
            queue = still
            active = [p for ps in self.packs.values() for p in ps if p.resident]
            if not active:
                if pending:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - (clock() - t0))))
                continue
            ts = clock()
            for p in active:
                p.step()
            torch.cuda.synchronize()
            ms = (clock() - ts) * 1e3
            rounds += 1
            if progress is not None and rounds % 50 == 0:
                progress("serving: round %d, %.1f s, %d resident, %d waiting"
                         % (rounds, clock() - t0, sum(len(p.resident) for p in active), len(queue) + len(pending)))
            for p in active:
                for rid, (req, x0) in list(p.resident.items()):
                    req.done_steps += 1
                    req.step_ms.append(ms)
                    cell_steps += req.cells()
                    if req.done_steps >= req.steps:
                        if keep_results:
                            req.result = p.extract(req, x0)
                        req.finished = clock() - t0
                        p.release(rid)
        wall = clock() - t0
        lat = np.array([m for r in requests for m in r.step_ms])
        return {"requests": len(requests), "rounds": rounds, "wall_s": wall,
                "cell_updates_per_s": cell_steps / wall if wall > 0 else 0.0,
                "step_latency_ms_p50": float(np.percentile(lat, 50)) if lat.size else None,
                "step_latency_ms_p99": float(np.percentile(lat, 99)) if lat.size else None,
                "packs": {"%s/%d" % k: len(v) for k, v in self.packs.items()}}


def warm_up(device="cuda"):
    """Short cases of each class through a throw-away server: three resident
    at once per class, then three more admitted while those run, and released
    in turn.  The first launch of each kernel a step uses (the library's, and
    torch's concatenations and index copies of the ghost and halo refresh,
    which differ with one and with several resident cases) loads its code:
    10-60 ms that a server pays once at start-up, not in a request's step."""
    reqs, rid = [], 0
    for arrival, steps in ((0.0, 4), (0.02, 3)):
        for kind, n in (("tube1", 256), ("tube5", 256), ("vortex", 64)):
            for k in range(3):
                reqs.append(Request(rid, kind, n + 32*k, steps + k, arrival))
                rid += 1
    Server(budget_bytes=1 << 30, tube_capacity=1 << 12, vortex_columns=512, device=device).run(reqs)
    if device == "cuda":
        torch.cuda.synchronize()


def poisson_requests(n, rate, seed=0, steps=(20, 60), vortex_sizes=(64, 128, 256)):
    """Seeded Poisson stream: shock tubes (N ~ U{256..16384}, pressure ratio
    ~ U[2, 20], 1 or 5 species) and 2-D vortices (N from vortex_sizes)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        kind = KINDS[rng.integers(0, 3)]
        if kind == "vortex":
            size = int(rng.choice(vortex_sizes))
        else:
            size = int(rng.integers(256, 16385))
        out.append(Request(i, kind, size, int(rng.integers(steps[0], steps[1] + 1)), t,
                           float(rng.uniform(2.0, 20.0))))
    return out
