"""Precision stress on the sports-wedge setup (1024x512x512, u = 0.1): the same setup on a 64^3
crop, GPU storage FP32 / FP16 / BF16 against the FP64 oracle with NO storage rounding, after 100
steps — the error each storage format adds to the populations and to C_D.

Bars: populations (max |f_gpu - f_64| / |f_64|) FP32 1e-5, FP16 2e-3, BF16 1e-2; C_D relative
error FP32 1e-4, FP16 1e-2, BF16 5e-2 (the force sums deviations over the wedge's wall links, where
the storage error does not average out). The measured errors are printed."""
import numpy as np
import pytest

from tests.cases import SEED, perturbed_state

pytestmark = pytest.mark.gpu

CROP = (64, 64, 64)
POP_TOL = {"f32": 1e-5, "f16": 2e-3, "bf16": 1e-2}
CD_TOL = {"f32": 1e-4, "f16": 1e-2, "bf16": 5e-2}


def crop_body():
    from windtunnel_vis_amd import geometry as G
    s = 64 / 512  # the crop's height over the full grid's
    return G.sports_wedge(CROP, length=300 * s, width=90 * s, height=45 * s, corner=10 * s, ride=1)


def crop_params(prec):
    return dict(u=0.1, nu=6e-4, cs=0.1, ground="moving", sides="freeslip", spongeNu=0.04, spongeStart=0.84, spongeIn=0.03,
                tauWall=0.53, interp=True, collision="regularized", emaAlpha=0.02, precision=prec)


@pytest.fixture(scope="module")
def oracle_run():
    """FP64 oracle (no storage rounding) on the crop, geometry voxelized by the GPU path."""
    from oracle.oracle import OracleSolver
    from windtunnel_vis_amd import SolverGPU
    from windtunnel_vis_amd.solver import D3Q19
    body = crop_body()
    g = SolverGPU(CROP, crop_params("f32"))
    g.setWheels(body.wheels)
    info = g.voxelizeSdf(body.prims())
    flags, sdf = g.readFlags(), g.readSdf()
    g.destroy()
    f0 = perturbed_state(D3Q19, CROP, 0.1, flags, seed=SEED, amp_u=0.01).astype(np.float32)
    o = OracleSolver(CROP, crop_params("f32"), storage=None)
    o.setWheels(body.wheels)
    o.uploadFlags(flags)
    o.uploadSdf(sdf)
    o.uploadPopulations(f0.astype(np.float64))
    st = o.stepAndSample(100)
    return dict(body=body, info=info, flags=flags, sdf=sdf, f0=f0, f=o.readPopulations(), fx=st.fx)


@pytest.mark.parametrize("prec", ["f32", "f16", "bf16"])
def test_crop_storage_error_vs_fp64(oracle_run, prec):
    from windtunnel_vis_amd import SolverGPU
    r = oracle_run
    g = SolverGPU(CROP, crop_params(prec))
    g.setWheels(r["body"].wheels)
    g.uploadFlags(r["flags"])
    g.uploadSdf(r["sdf"])
    g.uploadPopulations(r["f0"])
    st = g.stepAndSample(100)
    f = g.readPopulations().astype(np.float64)
    fo = r["f"]
    pop_err = float(np.max(np.abs(f - fo) / np.abs(fo)))
    cd_err = abs(st.fx - r["fx"]) / abs(r["fx"])
    print("crop %s: population error %.3e, C_D error %.3e (Fx %.6e vs %.6e)" % (prec, pop_err, cd_err, st.fx, r["fx"]))
    assert not st.unstable
    assert pop_err < POP_TOL[prec]
    assert cd_err < CD_TOL[prec]
    g.destroy()
