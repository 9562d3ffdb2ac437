"""Edge cases through the C ABI on the GPU: arguments the library must refuse (checked on the host,
before any launch), empty inputs, the half-width fallback for odd nx, and the smallest grids
against the oracle."""
import numpy as np
import pytest

from tests.cases import SEED, lattice_for, perturbed_state

pytestmark = pytest.mark.gpu

P3 = dict(u=0.05, nu=0.02, cs=0.1, ground="moving", sides="freeslip", spongeNu=0.03, spongeStart=0.8, spongeIn=0.0,
          tauWall=0.53, interp=True, collision="regularized", emaAlpha=0.1)


def test_refused_arguments():
    from windtunnel_vis_amd import SolverGPU
    from windtunnel_vis_amd._native import WtError
    with pytest.raises(WtError):
        SolverGPU((1, 8, 8), P3)  # nx >= 2
    with pytest.raises(WtError):
        SolverGPU((2048, 1024, 1024), P3)  # more than 2^31 cells
    s = SolverGPU((8, 6, 4), P3)
    with pytest.raises(WtError):
        s.step(-1)
    with pytest.raises(WtError):
        s.stepAndSample(-3)
    with pytest.raises(WtError):
        s.streamlines(dict(a=(1, 1, 1), b=(1, 4, 1), d=(0, 0, 0)), 2, 4, h=0.0)
    st = s.stepAndSample(0)  # no steps: the initial state's statistics
    assert st.step == 0 and not st.unstable
    s.destroy()


def test_empty_inputs():
    from windtunnel_vis_amd import SolverGPU
    s = SolverGPU((16, 8, 6), P3)
    flags = np.zeros(16 * 8 * 6, np.uint32)
    flags[200:210] = 1
    s.uploadFlags(flags)
    info = s.voxelizeMesh(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32), np.eye(4, dtype=np.float32).reshape(-1))
    assert info["solidCells"] == 0 and info["frontal"] == 0
    assert not s.readFlags().any()  # an empty mesh clears the geometry (VoxelizerGPU.ts:54-58)
    vel, mean = s.probe(np.zeros((0, 3)))
    assert vel.shape == (0, 4) and mean.shape == (0, 4)
    assert s.streamlines(dict(a=(1, 1, 1), b=(1, 4, 1), d=(0, 0, 0)), 0, 4).shape == (0, 4, 4)
    s.setWheels([])
    assert np.isfinite(s.stepAndSample(10).fx)
    s.destroy()


def test_odd_nx_falls_back_to_fp32():
    from windtunnel_vis_amd import SolverGPU
    s = SolverGPU((9, 6, 4), dict(P3, precision="f16"))
    assert s.precision == "f32"  # pairs of cells per 32-bit word need an even nx (SolverGPU.ts:69)
    s.destroy()


@pytest.mark.parametrize("dims,prec", [((2, 3, 2), "f32"), ((2, 3, 2), "f16"), ((4, 2, 1), "f32"), ((3, 2, 1), "f32")])
def test_smallest_grids_match_the_oracle(dims, prec):
    from oracle.oracle import OracleSolver
    from windtunnel_vis_amd import SolverGPU
    p = dict(P3, precision=prec, ground="noslip" if dims[2] == 1 else "moving")
    g = SolverGPU(dims, p)
    o = OracleSolver(dims, p, storage=None if prec == "f32" else prec)
    n = dims[0] * dims[1] * dims[2]
    flags = np.zeros(n, np.uint32)
    f0 = perturbed_state(lattice_for(dims), dims, p["u"], flags, seed=SEED).astype(np.float32)
    for s in (g, o):
        s.uploadFlags(flags)
    g.uploadPopulations(f0)
    o.uploadPopulations(f0.astype(np.float64))
    g.step(50)
    o.step(50)
    a, b = g.readPopulations().astype(np.float64), o.readPopulations()
    assert np.all(np.isfinite(a))
    err = np.max(np.abs(a - b) / np.abs(b))
    print(dims, prec, "max rel err %.3e" % err)
    assert err < (1e-5 if prec == "f32" else 2e-3)
    g.destroy()
