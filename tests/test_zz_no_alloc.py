"""htr_max_spectral_radius_noalloc (the entry point that allocates nothing)
against htr_max_spectral_radius through the C ABI: two calls in a row, the
second over a smaller rect whose maximum is below the first's (stale partial
maxima of the first call must not reach it), then an empty rect."""
import ctypes

import numpy as np
import pytest

from hotpath import air_state, complete_state, gpu_mix, to_gpu

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def device():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import htr_amd
    htr_amd.init(0)
    yield


def noalloc(fs, mix, grid, out):
    import htr_amd as H
    s = torch.cuda.current_stream()
    rc = H.lib().htr_max_spectral_radius_noalloc(ctypes.byref(grid), ctypes.byref(fs.c()), mix.h,
                                                 ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(s.cuda_stream))
    assert rc == 0, H.lib().htr_last_error()
    torch.cuda.synchronize()
    return float(out.item())


def test_noalloc_radius_equals_the_allocating_entry_point():
    import htr_amd as H
    n = (24, 24, 24)
    st, meta = air_state(24, seed=5)
    st = dict(st)
    v = st["velocity"].copy()
    v[:, 20, 20, 20] *= 6.0                  # the largest radius, outside the second rect
    st["velocity"] = v
    st, om = complete_state(st, meta, n, sensors=False)
    fs = to_gpu(st, 5, n)
    gm = gpu_mix(meta)
    out = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    full = fs.grid((0, 0, 0), (23, 23, 23))
    r1 = H.calculate_max_spectral_radius(fs, gm, grid=full)
    g1 = noalloc(fs, gm, full, out)
    print("full box: noalloc %.17g, allocating entry point %.17g" % (g1, r1), flush=True)
    assert g1 == r1
    out.fill_(-1.0)
    sub = fs.grid((2, 3, 4), (9, 8, 7))
    r2 = H.calculate_max_spectral_radius(fs, gm, grid=sub)
    g2 = noalloc(fs, gm, sub, out)
    print("sub-rect: noalloc %.17g, allocating entry point %.17g" % (g2, r2), flush=True)
    assert g2 == r2 and r2 < r1
    out.fill_(-1.0)
    empty = fs.grid((5, 5, 5), (4, 9, 9))
    assert noalloc(fs, gm, empty, out) == 0.0 == H.calculate_max_spectral_radius(fs, gm, grid=empty)
