"""Probes, streamlines and particle tracers (render/tracers.ts) on the GPU against the oracle, on a
velocity field the GPU's own macro pass produced (a smooth linear flow around a solid block)."""
import numpy as np
import pytest

from tests.cases import feq
from windtunnel_vis_amd.solver import D2Q9, D3Q19

pytestmark = pytest.mark.gpu


def _setup(dims):
    from windtunnel_vis_amd import SolverGPU
    nx, ny, nz = dims
    lat = D3Q19 if nz > 1 else D2Q9
    s = SolverGPU(dims, dict(u=0.04, nu=0.02, cs=0.0, emaAlpha=0.5, precision="f32"))
    z, y, x = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    x, y, z = x.reshape(-1), y.reshape(-1), z.reshape(-1)
    flags = ((np.abs(x - 0.6 * nx) < 2) & (np.abs(y - ny / 2) < 2) & (np.abs(z - nz / 2) < 2 if nz > 1 else True)).astype(np.uint32)
    u = np.stack([0.04 + 0.0005 * (y - ny / 2), 0.0003 * (x - nx / 2), (0.0002 * (z - nz / 2)) if nz > 1 else 0 * x])
    u[:, flags != 0] = 0.0
    s.uploadFlags(flags)
    s.uploadPopulations(feq(lat, np.ones(len(x)), u).astype(np.float32))
    s.stepAndSample(0)  # macro pass: velocity record from these populations
    s.derive()
    vel, _, _ = s.readMacro()
    _, _, mean = s.readDerived()
    return s, flags != 0, vel, mean


@pytest.mark.parametrize("dims", [(24, 16, 12), (40, 20, 1)])
def test_probes_match_oracle(dims):
    from oracle import fields_oracle as FO
    s, solid, vel, mean = _setup(dims)
    rng = np.random.default_rng(11)
    p = rng.uniform(-0.5, 1.0, (200, 3)) * np.array(dims, dtype=float)[None, :] * np.array([1, 1, 1 if dims[2] > 1 else 0])
    v, m = s.probe(p)
    wv, wm = FO.sample(vel, dims, p), FO.sample(mean, dims, p)
    if dims[2] == 1:
        wv[:, 2] = wm[:, 2] = 0
    np.testing.assert_allclose(v, wv, atol=2e-7, rtol=1e-6)
    np.testing.assert_allclose(m, wm, atol=2e-7, rtol=1e-6)
    s.destroy()


@pytest.mark.parametrize("dims", [(24, 16, 12), (40, 20, 1)])
def test_streamlines_match_oracle(dims):
    from oracle import fields_oracle as FO
    s, solid, vel, mean = _setup(dims)
    rake = dict(a=(1.5, 2.0, 3.0), b=(1.5, dims[1] - 2.0, 3.0), d=(0.0, 0.0, 1.5))
    for use_mean, field in ((False, vel), (True, mean)):
        got = s.streamlines(rake, seeds=12, points=48, h=0.5, perLine=6, useMean=use_mean)
        want = FO.streamlines(field, solid, dims, rake, 12, 6, 48, 0.5, 0.04)
        assert np.array_equal(got[:, :, 3] < 0, want[:, :, 3] < 0)
        np.testing.assert_allclose(got, want, atol=2e-4)
        assert (got[:, :, 3] < 0).any() and (got[:, 0, 3] > 0).all()  # some lines stop at the block or the outlet
    s.destroy()


@pytest.mark.parametrize("emitter", ["inlet", "rake"])
def test_particles_match_oracle(emitter):
    from oracle import fields_oracle as FO
    from windtunnel_vis_amd.solver import Tracers
    dims = (24, 16, 12)
    s, solid, vel, mean = _setup(dims)
    rake = dict(a=(2.0, 3.0, 6.0), b=(2.0, 13.0, 6.0), d=(0.0, 0.0, 2.0))
    # maxAge 2: particles older than 8 frames respawn, so the 12 frames exercise every respawn path
    t = Tracers(s, count=300, trail=6, emitter=emitter, nozzles=10, lines=2, maxAge=2.0, rake=rake)
    o = dict(count=300, trail=6, emitter=0 if emitter == "inlet" else 1, nozzles=10, lines=2, max_age=2.0, rake=rake)
    st = np.zeros((300, 4))
    hist = np.zeros((6, 300, 4))
    head = 0
    for frame in range(12):
        fill = frame == 0
        t.update(20.0, advance=True, fill=fill, seed=frame)
        if not fill:
            head = (head + 1) % 6
        FO.particles_update(st, hist, head, vel, solid, dims, o, 20.0, frame, fill, True)
    gst, ghist, ghead = t.read()
    assert ghead == head
    np.testing.assert_allclose(gst, st, atol=2e-4)
    np.testing.assert_allclose(ghist, hist, atol=2e-4)
    assert (gst[:, 3] < 4).any()  # respawns happened (outlet, block or age)
    t.destroy()
    s.destroy()
