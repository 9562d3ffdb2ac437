"""Derived fields (wgslLbm.ts:527-592) and the strided export (SolverGPU.ts:374-425) on the GPU
against the oracle, on the fields the GPU's own macro pass produced.

Tolerances: the GPU computes in FP32 from the half velocity records and stores halves; the oracle
computes in FP64 and is compared after half rounding. Allowed: 2 half ulps of the value plus the
half subnormal step (2^-24); for TI also the FP32 cancellation in meanB - |meanA|^2. The mean copy,
the solid flag and every export are bit-exact."""
import numpy as np
import pytest

from tests.cases import SEED, lattice_for, make_case, perturbed_state

pytestmark = pytest.mark.gpu

H_ULP = 2.0 ** -10
H_SUB = 2.0 ** -24


def _run(name, prec, steps=60):
    from windtunnel_vis_amd import SolverGPU
    dims, params, flags, sdf, wheels = make_case(name)
    lat = lattice_for(dims)
    g = SolverGPU(dims, dict(params, precision=prec))
    g.setWheels(wheels)
    g.uploadFlags(flags)
    g.uploadSdf(sdf)
    g.uploadPopulations(perturbed_state(lat, dims, params["u"], flags, seed=SEED).astype(np.float32))
    for _ in range(steps // 20):  # several samples so the EMA statistics move
        g.stepAndSample(20)
    g.derive()
    return g, dims, params, flags


def _close(gpu, want, extra=0.0):
    w = np.asarray(want, dtype=np.float64)
    err = np.abs(gpu.astype(np.float64) - w)
    tol = 2 * H_ULP * np.abs(w) + H_SUB + extra
    bad = err > tol
    assert not bad.any(), (int(bad.sum()), float(err.max()), np.argwhere(bad)[:5].tolist())


@pytest.mark.parametrize("name,prec", [("d3_full", "f32"), ("d3_full", "f16"), ("d3_odd_f32", "f32"), ("d2_cyl", "f16")])
def test_derived_matches_oracle(name, prec):
    from oracle import fields_oracle as FO
    g, dims, params, flags = _run(name, prec)
    vel, ma, mb = g.readMacro()
    vort, stat, mean = g.readDerived()
    solid = flags != 0
    w_vort, w_stat, w_mean = FO.derived(vel, solid, dims, ma, mb, max(params["u"], 1e-4))
    print(name, prec, "max |Q|", float(np.abs(w_vort[:, 3]).max()), "max TI", float(w_stat[:, 0].max()))
    assert np.abs(w_vort[:, :3]).max() > 1e-3 and np.abs(w_vort[:, 3]).max() > 1e-6  # the flow is not trivial
    _close(vort, w_vort)
    kerr = 4 * 2.0 ** -24 * (mb.astype(np.float64) + (ma[:, :3].astype(np.float64) ** 2).sum(1))
    _close(stat[:, 0], w_stat[:, 0], extra=np.sqrt(kerr / 3) / max(params["u"], 1e-4))
    np.testing.assert_array_equal(stat[:, 1:], w_stat[:, 1:])
    np.testing.assert_array_equal(mean, FO.half(w_mean).astype(np.float32))
    g.destroy()


@pytest.mark.parametrize("stride", [1, 2, 3, 5])
def test_export_samples_fields(stride):
    from oracle import fields_oracle as FO
    g, dims, params, flags = _run("d3_full", "f16", steps=40)
    vel, _, _ = g.readMacro()
    vort, _, mean = g.readDerived()
    f = g.exportFields(stride)
    assert f.dims == FO.export_dims(dims, stride) and f.stride == stride
    np.testing.assert_array_equal(f.vel, FO.export_sample(vel, dims, stride))
    np.testing.assert_array_equal(f.mean, FO.export_sample(mean, dims, stride))
    np.testing.assert_array_equal(f.vort, FO.export_sample(vort, dims, stride))
    g.destroy()


def test_export_before_any_derive_is_zero_and_stride_checked():
    from windtunnel_vis_amd import SolverGPU
    from windtunnel_vis_amd._native import WtError
    g = SolverGPU((16, 8, 8), dict(u=0.05, nu=0.02))
    g.stepAndSample(2)
    f = g.exportFields(4)
    assert f.dims == (4, 2, 2)
    assert not f.mean.any() and not f.vort.any()
    assert np.allclose(f.vel[:, 0], 0.05, atol=1e-3)
    with pytest.raises(WtError):
        g.exportFields(0)
    g.destroy()


def test_vtk_from_gpu_export():
    from windtunnel_vis_amd import export as E
    g, dims, params, flags = _run("d3_full", "f16", steps=40)
    st = g.stepAndSample(0)
    g.derive()
    f = g.exportFields(E.stride_for(g.n, 500))
    data = E.vtk(f, dx=0.01, U=params["u"], L=10.0, rhoRef=st.rhoRef, title="d3_full")
    hdr, arr = E.read_vtk(data)
    assert hdr["dims"] == f.dims
    np.testing.assert_array_equal(arr["velocity"], (f.vel[:, :3].astype(np.float64) / params["u"]).astype(np.float32))
    g.destroy()
