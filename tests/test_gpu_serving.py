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
