"""The plain D2Q9 BGK cylinder workload (512 x 128, FP32, Re = 100, seeded perturbation) on the
GPU: over the developed-shedding window (steps 20k-40k) the Strouhal number must fall in the
reference's band (validation.ts:125, St in [0.16, 0.20] over >= 4 cycles) and C_D near the Re = 100
value for a staircase cylinder at 16 % blockage (1.3-1.8). The configuration's own 20k-step window
must run stably and show the onset of shedding."""
import pytest

pytestmark = pytest.mark.gpu


def test_cylinder_512x128_strouhal_and_drag():
    from windtunnel_vis_amd import SolverGPU
    from windtunnel_vis_amd.validation import cylinder_512x128
    print("cylinder 512x128 D2Q9 FP32, 40k steps", flush=True)
    r = cylinder_512x128(lambda d, p: SolverGPU(d, p))
    print(r, flush=True)
    assert not r["unstable"]
    dev = r["developed"]
    assert dev["cycles"] >= 4
    assert 0.16 <= dev["St"] <= 0.20
    assert 1.3 <= dev["cd"] <= 1.8
    w = r["window_20k"]
    assert w["cl_amp"] > 0.05 and 1.3 <= w["cd"] <= 1.8
