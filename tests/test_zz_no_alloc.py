"""wt_mlp_infer_noalloc (the learned closure's stand-alone inference with its scratch in a static
device array, for callers that cannot allocate) against wt_mlp_infer through the C ABI: bit for bit
on two calls in a row, the second smaller, then the size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(lib, fn, m, raw):
    from windtunnel_vis_amd import _native as N
    out = np.full(len(raw), np.nan, np.float32)
    N.check(getattr(lib, fn)(ctypes.byref(m), N.ptr(raw), len(raw), N.ptr(out)))
    return out


def test_noalloc_inference_equals_allocating_inference():
    from windtunnel_vis_amd import _native as N
    from windtunnel_vis_amd import closure
    lib = N.load()
    rng = np.random.default_rng(31)
    net = closure.Network.initial(seed=4, nu_scale=2e-3)
    net.mu = rng.normal(size=8).astype(np.float32)
    net.sigma = np.full(8, 2.5, np.float32)
    m = net.to_struct()
    # the maximum size first, then a smaller ragged one: the second call must not see the first's data
    for n in (65536, 1000, 129):
        raw = (rng.normal(size=(n, 8)) * 3).astype(np.float32)
        want = _call(lib, "wt_mlp_infer", m, raw)
        got = _call(lib, "wt_mlp_infer_noalloc", m, raw)
        print(f"noalloc inference n={n}: max |diff| {np.abs(got - want).max():.3e}", flush=True)
        assert np.isfinite(want).all()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # a different network on the same scratch
    net2 = closure.Network.initial(seed=9, nu_scale=5e-4)
    m2 = net2.to_struct()
    raw = (rng.normal(size=(300, 8)) * 3).astype(np.float32)
    assert np.array_equal(_call(lib, "wt_mlp_infer_noalloc", m2, raw).view(np.uint32),
                          _call(lib, "wt_mlp_infer", m2, raw).view(np.uint32))


def test_noalloc_inference_rejects_oversize_input():
    from windtunnel_vis_amd import _native as N
    from windtunnel_vis_amd import closure
    lib = N.load()
    m = closure.Network.initial(seed=1).to_struct()
    raw = np.zeros((65537, 8), np.float32)
    out = np.zeros(65537, np.float32)
    rc = lib.wt_mlp_infer_noalloc(ctypes.byref(m), N.ptr(raw), len(raw), N.ptr(out))
    assert rc == N.WT_E_ARG
    assert b"WT_MLP_NOALLOC_MAX" in lib.wt_last_error()
    assert lib.wt_mlp_infer_noalloc(ctypes.byref(m), N.ptr(raw), 0, N.ptr(out)) == N.WT_OK
