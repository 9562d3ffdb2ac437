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


@pytest.mark.parametrize("prec", ["f16", "f32"])
def test_batch_matches_oracle(prec):
    """Each member of a batched run against the FP64 oracle stepped on its own (not only against the
    single-domain kernel): 60 steps in graph chunks of 20, the FP16 / FP32 bars."""
    from oracle.oracle import OracleSolver
    from windtunnel_vis_amd import BatchGPU
    _, params, _, _, _ = make_case("d3_full")
    A = _make(prec)
    O = []
    for k, dims in enumerate(DIMS):
        flags, sdf, wheels = box_body_3d(dims)
        o = OracleSolver(dims, dict(params, precision=prec), storage="f16" if prec == "f16" else None)
        o.setWheels(wheels)
        o.uploadFlags(flags)
        o.uploadSdf(sdf)
        o.uploadPopulations(perturbed_state(lattice_for(dims), dims, params["u"], flags, seed=SEED + k).astype(np.float32)
                            .astype(np.float64))
        O.append(o)
    batch = BatchGPU(A, chunk=20, graph=True)
    batch.run(60)
    st = batch.stats()
    tol = 2e-3 if prec == "f16" else 1e-5
    for k, (a, o) in enumerate(zip(A, O)):
        so = o.stepAndSample(60)
        fa, fo = a.readPopulations().astype(np.float64), o.readPopulations()
        err = np.max(np.abs(fa - fo) / np.abs(fo))
        print(f"batch member {k} {prec}: max rel err vs oracle {err:.3e}", flush=True)
        assert err < tol
        assert abs(st[k].fx - so.fx) <= (5e-3 if prec == "f16" else 1e-4) * max(1.0, abs(so.fx))
    batch.destroy()
    for s in A + O:
        s.destroy()


def test_batch_refuses_a_member_that_left_the_kernel_variant():
    """Tuning, collision or an attached network changed on a member after the batch was made: the
    batch refuses to run (WT_E_STATE) instead of stepping the member with the wrong kernel; undoing
    the change makes it run again."""
    from windtunnel_vis_amd import BatchGPU, closure
    from windtunnel_vis_amd._native import WT_E_STATE, WtError
    A, B = _make("f32"), _make("f32")
    batch = BatchGPU(A, chunk=4, graph=True)
    batch.run(4)
    for s in B:
        s.step(4)
    changes = [
        (lambda: A[1].setTuning(1, 1), lambda: A[1].setTuning(1, 2)),
        (lambda: A[2].setParams({"collision": "bgk"}), lambda: A[2].setParams({"collision": "regularized"})),
    ]
    net = closure.Network(nu_scale=1e-3, smag_floor=1.0)
    changes.append((lambda: closure.attach(A[0], net), lambda: A[0].setParams({"closure": "smagorinsky"})))
    for change, undo in changes:
        change()
        with pytest.raises(WtError) as e:
            batch.run(4)
        assert e.value.code == WT_E_STATE
        undo()
        batch.run(4)
        for s in B:
            s.step(4)
    _same(A, B)
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
