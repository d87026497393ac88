"""Continuous-batching request loop: a case's result is the same alone and
in a mix of other cases that are admitted while it runs (GPU)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import htr_amd
    htr_amd.init(0)


def serve(reqs, **kw):
    from htr_amd import serving
    srv = serving.Server(**kw)
    stats = srv.run(reqs, keep_results=True)
    return stats


def test_alone_equals_mixed():
    from htr_amd.serving import Request
    mk = lambda: [Request(0, "tube5", 300, 6, 0.0, 10.0), Request(1, "vortex", 32, 6, 0.0)]
    alone = {}
    for r in mk():
        serve([r], tube_capacity=1024, vortex_columns=88)
        alone[r.id] = r.result
    mix = mk() + [Request(2, "tube5", 500, 3, 0.0, 4.0), Request(3, "tube5", 257, 4, 0.0, 18.0),
                  Request(4, "vortex", 32, 2, 0.0), Request(5, "vortex", 32, 3, 0.0),
                  Request(6, "tube1", 700, 5, 0.0, 7.0)]
    stats = serve(mix, tube_capacity=1024, vortex_columns=88, max_packs_per_class=1)
    print("mix stats", stats, flush=True)
    first = min(r.admitted for r in mix)
    late = [r.id for r in mix if r.admitted > first + 1e-9 and r.admitted_round > 0]
    print("admitted after the start:", late, flush=True)
    for rid in (0, 1):
        a, b = alone[rid], mix[rid].result
        assert a.shape == b.shape and np.all(np.isfinite(a))
        d = float(np.max(np.abs(a - b)))
        print("case %d: max |alone - mixed| = %.3e" % (rid, d), flush=True)
        assert d == 0.0
    # one pack per class of fixed capacity: some cases had to wait for a free
    # segment and were admitted while others were running
    assert late, "no case was admitted into a running pack"
    assert stats["rounds"] > 6


def oracle_case(req):
    """The case alone on its own box, advanced by the oracle with the same
    step sequence (sensors, RHS at deltaTime dt RK_C[stage][3], SSP-RK3
    update, primitives, properties, boundary values): the vortex on a
    periodic n x n box, the tube with its two Dirichlet ghost cells."""
    from oracle import oracle as O
    from htr_amd import serving
    from hotpath import oracle_fields, oracle_mix
    st, meta = req.state()
    om = oracle_mix(meta)
    n = (st["rho"].shape[-1] if "rho" in st else st["pressure"].shape[-1], st["pressure"].shape[-2], 1)
    g, f = oracle_fields(st, om.nSpec, n)
    O.update_properties(g, f, om)
    O.update_conserved(g, f, om)
    dt = serving.DT[req.kind]
    dirs = (0, 1) if req.kind == "vortex" else (0,)
    keep = ("Conserved", "temperature", "pressure", "MolarFracs", "velocity", "rho", "mu", "lam", "Di", "SoS",
            "MassFracs")
    ghost = None
    if req.kind != "vortex":
        ghost = {k: f[k][..., [0, -1]].copy() for k in keep}
    RK = ((1.0, 0.0, 1.0), (0.75, 0.25, 0.25), (1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0))
    for _ in range(req.steps):
        f["Conserved_old"] = f["Conserved"]
        for stage in range(3):
            O.ducros_sensor(g, f, 1.0)
            for d in dirs:
                O.shock_sensor(g, f, d, om)
            f["Conserved_t"] = 0.0
            if om.nSpec > 1:
                O.chem_source(g, f, om)
            for d in reversed(dirs):
                O.euler_flux(g, f, d, O.SCHEME_HYBRID, RK[stage][0], RK[stage][1], dt * RK[stage][2], om)
            for d in reversed(dirs):
                O.diffusion_flux(g, f, d, om)
            O.rk_update(g, f, om.nSpec + 4, stage, dt)
            O.update_primitive(g, f, om)
            O.update_properties(g, f, om)
            if ghost is not None:
                for k in keep:
                    f[k][..., [0, -1]] = ghost[k]
    c = f["Conserved"]
    return c if req.kind == "vortex" else c[..., 1:-1]


def test_served_cases_match_the_oracle():
    """A 5-species shock tube and a 64^2 vortex served together (with other
    cases admitted around them) equal the oracle's run of each case on its
    own box: the halos and ghosts of a pack reproduce the case's boundaries."""
    from htr_amd.serving import Request
    reqs = [Request(0, "tube5", 300, 3, 0.0, 10.0), Request(1, "vortex", 64, 3, 0.0),
            Request(2, "tube5", 257, 2, 0.0, 4.0), Request(3, "vortex", 64, 2, 0.0),
            Request(4, "tube1", 400, 3, 0.0, 7.0)]
    serve(reqs, tube_capacity=2048, vortex_columns=200)
    for r in reqs[:2] + reqs[4:]:
        ref = oracle_case(r)
        got = r.result
        assert got.shape == ref.shape, (got.shape, ref.shape)
        # components that vanish identically (the transverse momenta of a
        # tube) must vanish on both sides: normalised by 1 there
        scale = np.max(np.abs(ref.reshape(ref.shape[0], -1)), axis=1)
        err = np.max(np.abs(got - ref).reshape(ref.shape[0], -1), axis=1) / np.where(scale > 0, scale, 1.0)
        print("case %d (%s %d): max normalised error vs oracle %.3e" % (r.id, r.kind, r.n, float(err.max())),
              flush=True)
        assert np.all(np.isfinite(got)) and float(err.max()) <= 1e-12


def test_large_vortex_stays_finite():
    """A 1024^2 vortex (the largest of the serving configuration) for two
    steps: finite, mass conserved to round-off."""
    from htr_amd.serving import Request
    r = Request(0, "vortex", 1024, 2, 0.0)
    stats = serve([r], vortex_columns=1024 + 12)
    st, meta = r.state()
    out = r.result
    print("1024^2 vortex: %s, step ms %s" % (stats["step_latency_ms_p50"], r.step_ms), flush=True)
    assert np.all(np.isfinite(out))
    from htr_amd import FieldSet
    import htr_amd as H
    from hotpath import gpu_mix
    m0 = None
    fs = FieldSet((1024, 1024, 1), 1)
    for k, v in st.items():
        if k in fs.t:
            fs[k] = v
    H.update_properties_from_primitive(fs, gpu_mix(meta))
    H.update_conserved_from_primitive(fs, gpu_mix(meta))
    m0 = float(fs["Conserved"][0].sum())
    m1 = float(out[0].sum())
    assert abs(m1 - m0) <= 1e-11 * abs(m0)
