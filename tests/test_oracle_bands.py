"""The reference's validation bands (validation.ts) reproduced by the CPU oracle at reduced scale
(valscale 0.5, as the reference's CI runs them, scripts/harness.mjs:11-15)."""
import pytest

from windtunnel_vis_amd import validation as V


def _make(d, p):
    from oracle.oracle import OracleSolver
    return OracleSolver(d, p)


def test_freestream_band_oracle():
    r = V.freestream(_make)
    assert r["passed"], r


@pytest.mark.slow
def test_cylinder_band_oracle():
    r = V.cylinder(_make, scale=0.5)
    print(r)
    assert r["passed"], r
