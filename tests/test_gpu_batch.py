"""Batched sweeps on the GPU: one launch per step for many uneven domains (descriptor table),
replayed hipGraph chunks and direct launches must give exactly what stepping each domain on its
own gives (same kernels' arithmetic, every item covered once)."""
import numpy as np
import pytest

from tests.cases import SEED, box_body_3d, lattice_for, make_case, perturbed_state

pytestmark = pytest.mark.gpu

DIMS = [(32, 16, 12), (48, 12, 16), (24, 20, 10)]


def _make(prec):
    from windtunnel_vis_amd import SolverGPU
    _, params, _, _, _ = make_case("d3_full")
    out = []
    for k, dims in enumerate(DIMS):
        flags, sdf, wheels = box_body_3d(dims)
        s = SolverGPU(dims, dict(params, precision=prec))
        s.setWheels(wheels)
        s.uploadFlags(flags)
        s.uploadSdf(sdf)
        s.uploadPopulations(perturbed_state(lattice_for(dims), dims, params["u"], flags, seed=SEED + k).astype(np.float32))
        out.append(s)
    return out


def _same(a, b):
    for k, (x, y) in enumerate(zip(a, b)):
        assert np.array_equal(x.readPopulations().view(np.uint32), y.readPopulations().view(np.uint32)), k


@pytest.mark.parametrize("prec", ["f16", "f32"])
def test_batch_matches_individual_steps(prec):
    from windtunnel_vis_amd import BatchGPU
    A, B = _make(prec), _make(prec)
    batch = BatchGPU(A, chunk=4, graph=True)
    print("batch %s: graph chunk 4, 10 steps" % prec, flush=True)
    batch.run(10)  # two replays + two direct launches
    for s in B:
        s.step(10)
    _same(A, B)
    # members stepped on their own in between (eager work between replays), then the batch from
    # the other parity: a second captured chunk
    for s in A + B:
        s.step(1)
    batch.run(9)
    for s in B:
        s.step(9)
    _same(A, B)
    batch.run(8)  # replays of the first chunk again
    for s in B:
        s.step(8)
    _same(A, B)
    st = batch.stats()
    for s, a in zip(B, st):
        b = s.stepAndSample(0)
        assert (a.fx, a.fy, a.fz, a.maxU, a.rhoRef, a.step) == (b.fx, b.fy, b.fz, b.maxU, b.rhoRef, b.step)
    batch.destroy()
    for s in A + B:
        s.destroy()


def test_batch_direct_launches_and_validation():
    from windtunnel_vis_amd import BatchGPU, SolverGPU
    from windtunnel_vis_amd._native import WtError
    A, B = _make("f16"), _make("f16")
    batch = BatchGPU(A, chunk=6, graph=False)
    batch.run(7)
    for s in B:
        s.step(7)
    _same(A, B)
    batch.destroy()
    # members must share the storage format and parity
    _, params, _, _, _ = make_case("d3_full")
    odd = SolverGPU((32, 16, 12), dict(params, precision="f32"))
    with pytest.raises(WtError):
        BatchGPU([A[0], odd])
    A[1].step(1)
    with pytest.raises(WtError):
        BatchGPU([A[0], A[1]])
    for s in A + B + [odd]:
        s.destroy()
