"""The reference's own validation bands (src/analysis/validation.ts), on the MI355X solver at the
reference's full case sizes (valscale 1) and storage formats."""
import numpy as np
import pytest

from windtunnel_vis_amd import validation as V

pytestmark = pytest.mark.gpu


def _make(prec):
    from windtunnel_vis_amd import SolverGPU
    return lambda d, p: SolverGPU(d, dict(p, precision=prec))


@pytest.mark.parametrize("prec", ["f32", "f16"])
def test_freestream(prec):
    r = V.freestream(_make(prec))
    print(r)
    assert r["passed"]


@pytest.mark.parametrize("prec", ["f32", "f16"])
def test_cylinder_strouhal(prec):
    r = V.cylinder(_make(prec), scale=1.0)
    print(r)
    assert r["passed"]


@pytest.mark.parametrize("prec", ["f32", "f16"])
def test_sphere_drag(prec):
    r = V.sphere(_make(prec), scale=1.0)
    print(r)
    assert r["passed"]


def _oracle_fixture(key):
    import json
    import os
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "validation_oracle.json")
    return json.load(open(p))[key]["metrics"]


@pytest.mark.parametrize("prec,tol_st,tol_cd", [("f32", 0.01, 0.005), ("f16", 0.02, 0.01)])
def test_cylinder_st_and_cd_against_fp64_oracle(prec, tol_st, tol_cd):
    """The cylinder case at the reference CI's size (valscale 0.5: D = 12, 300 x 120) against the FP64
    oracle's run of the same case (tests/golden/validation_oracle.json, make_validation_oracle.py):
    relative error of St and of the mean C_D."""
    ref = _oracle_fixture("cylinder_valscale_0.5")
    r = V.cylinder(_make(prec), scale=0.5)
    m = r["metrics"]
    e_st = abs(m["St"] - ref["St"]) / ref["St"]
    e_cd = abs(m["cd"] - ref["cd"]) / ref["cd"]
    print(f"cylinder {prec}: St {m['St']:.5f} (oracle {ref['St']:.5f}, err {e_st:.2e}), "
          f"C_D {m['cd']:.5f} (oracle {ref['cd']:.5f}, err {e_cd:.2e}), cycles {m['cycles']}", flush=True)
    assert r["passed"]
    assert e_st < tol_st and e_cd < tol_cd


@pytest.mark.parametrize("prec", ["f32", "f16"])
def test_ahmed_against_fp64_oracle(prec):
    """The Ahmed case at the reference CI's size (valscale 0.5, L = 32) against the FP64 oracle's run
    of it. The oracle (the reference's kernels restated) reads C_D near 0.78 there, far outside the
    reference's own band (0.285 +- 35 %): the band is not reachable by the reference's method at
    these resolutions. The GPU run must read the oracle's value within 3 standard errors of the two
    windows plus 3 %."""
    ref = _oracle_fixture("ahmed_valscale_0.5")
    r = V.ahmed(_make(prec), scale=0.5)
    m = r["metrics"]
    tol = 3 * (m["se"] ** 2 + ref["se"] ** 2) ** 0.5 + 0.03 * ref["cd"]
    print(f"ahmed L=32 {prec}: C_D {m['cd']:.4f} +- {m['se']:.4f} (oracle {ref['cd']:.4f} +- {ref['se']:.4f}), tol {tol:.4f}",
          flush=True)
    assert abs(m["cd"] - ref["cd"]) < tol


@pytest.mark.parametrize("prec", ["f32", "f16"])
def test_ahmed_drag(prec):
    """The Ahmed case at full size (valscale 1, L = 64 cells, Re_sim = 2e4) against the FP64 oracle's
    run of it (about 52 minutes on 5 cores): the oracle reads C_D 0.490 +- 0.019, so the reference's
    band (0.285 +- 35 %) is out of reach of its own method at this resolution too (its README notes
    that C_D reads high and publishes no Ahmed figure). The GPU run must be stable, converged and
    within 3 standard errors of the two windows plus 3 % of the oracle's value."""
    ref = _oracle_fixture("ahmed_valscale_1")
    r = V.ahmed(_make(prec), scale=1.0)
    m = r["metrics"]
    tol = 3 * (m["se"] ** 2 + ref["se"] ** 2) ** 0.5 + 0.03 * ref["cd"]
    print(f"ahmed L=64 {prec}: C_D {m['cd']:.4f} +- {m['se']:.4f} (oracle {ref['cd']:.4f} +- {ref['se']:.4f}), tol {tol:.4f}",
          flush=True)
    assert np.isfinite(m["cd"]) and m["se"] < 0.05
    assert abs(m["cd"] - ref["cd"]) < tol
