"""Launch configurations and the autotuner on the GPU.

Every item-run length and cells-per-thread choice must give bit-identical populations (the item
mapping covers each item exactly once; the arithmetic per cell does not depend on the launch), and
the autotuner must leave the solver's state as it found it.
"""
import numpy as np
import pytest

from tests.cases import SEED, box_body_3d, lattice_for, make_case, perturbed_state

pytestmark = pytest.mark.gpu


def _solver(name, prec, dims=None):
    from windtunnel_vis_amd import SolverGPU
    d0, params, flags, sdf, wheels = make_case(name)
    if dims is not None:
        flags, sdf, wheels = box_body_3d(dims)
    else:
        dims = d0
    g = SolverGPU(dims, dict(params, precision=prec))
    g.setWheels(wheels)
    g.uploadFlags(flags)
    g.uploadSdf(sdf)
    f0 = perturbed_state(lattice_for(dims), dims, params["u"], flags, seed=SEED).astype(np.float32)
    g.uploadPopulations(f0)
    return g, f0


def _chunks(g):
    nx, ny = g.dims["nx"], g.dims["ny"]
    wpr = -(-nx // 128)
    return [4, 1, 3, wpr, 8 * wpr, wpr * ny, 4 * wpr * ny, 0]


@pytest.mark.parametrize("name,prec,dims", [("d3_full", "f16", None), ("d3_full", "f32", None),
                                            ("d3_full", "bf16", (200, 9, 7)), ("d2_cyl", "f32", None)])
def test_launch_configurations_agree(name, prec, dims):
    g, f0 = _solver(name, prec, dims)
    refs = {}
    cpts = [2, 1] if g.precision == "f32" and g.dims["nx"] % 2 == 0 else [2]
    for cpt in cpts:
        g.setTuning(1, cpt)
        for ck in _chunks(g):
            g.setTuning(0, ck)
            g.uploadPopulations(f0)
            g.step(25)
            out = g.readPopulations()
            print("case %s %s cpt %d chunk %d" % (name, prec, cpt, ck), flush=True)
            if cpt not in refs:
                refs[cpt] = out
            else:  # same kernel, other item order: bit-identical
                assert np.array_equal(out.view(np.uint32), refs[cpt].view(np.uint32)), (cpt, ck)
    if len(refs) == 2:
        # cells per thread changes which cells take the wall path (other FMA contractions): FP32 rounding
        err = np.max(np.abs(refs[1].astype(np.float64) - refs[2]))
        assert err < 2e-6, err
    g.destroy()


def test_autotune_restores_state_and_applies_choice(tmp_path):
    from windtunnel_vis_amd import tuning
    g, f0 = _solver("d3_full", "f16")
    g.step(7)
    before = g.readPopulations()
    step0 = g.stepCount
    r = tuning.autotune(g, steps=5, cache=str(tmp_path / "t.json"))
    assert not r["cached"] and len(r["trials"]) >= 2
    assert g.stepCount == step0
    assert np.array_equal(g.readPopulations().view(np.uint32), before.view(np.uint32))
    assert g.getTuning(0) == r["chunk"] and g.getTuning(1) == r["cellsPerThread"]
    # a second call is served from the cache
    r2 = tuning.autotune(g, steps=5, cache=str(tmp_path / "t.json"))
    assert r2["cached"] and r2["chunk"] == r["chunk"]
    g.destroy()


def test_profile_modes_count_launches_and_leave_results_alone():
    g, f0 = _solver("d3_full", "f16")
    ref, _ = _solver("d3_full", "f16")
    ref.step(9)
    g.profileEnable("launch")
    g.step(3)
    ms_l, n_l, _ = g.profileRead()
    g.profileEnable("span")
    g.step(4)
    g.step(0)
    g.step(2)
    ms_s, n_s, _ = g.profileRead()
    g.profileEnable(False)
    assert (n_l, n_s) == (3, 6) and ms_l > 0 and ms_s > 0
    ms_0, n_0, _ = g.profileRead()
    assert (ms_0, n_0) == (0.0, 0)
    # timing changes no result: 9 profiled steps equal 9 plain ones
    assert np.array_equal(g.readPopulations().view(np.uint32), ref.readPopulations().view(np.uint32))
    with pytest.raises(ValueError):
        g.profileEnable("sometimes")
    g.destroy()
    ref.destroy()
