"""The reference's own validation bands (src/analysis/validation.ts), on the MI355X solver at the
reference's full case sizes (valscale 1) and storage formats."""
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
