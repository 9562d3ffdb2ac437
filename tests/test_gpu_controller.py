"""The run controller drives the MI355X solver: a sphere in a 96x48x48 tunnel, 120 frames of 20
steps with the start-up ramps, EMA schedule and watchdog active."""
import math

import pytest

pytestmark = pytest.mark.gpu


def test_controller_runs_sphere():
    from windtunnel_vis_amd import SolverGPU
    from windtunnel_vis_amd import geometry as G
    from windtunnel_vis_amd.controller import RunController
    dims = (96, 48, 48)
    s = SolverGPU(dims, dict(u=0.05, nu=0.01, cs=0.1, ground="freeslip", spongeNu=0.03, spongeStart=0.85, emaAlpha=0.02,
                             precision="f16"))
    body = G.sphere_body(dims, (30, 24, 24), 6)
    info = s.voxelizeSdf(body.prims())
    c = RunController(s, Lcells=12, frontal=info["frontal"], speed_kmh=120, reOverride=300, ground="freeslip")
    for _ in range(120):
        c.frame(20)
    r = c.readout()
    print({k: r[k] for k in ("cd", "cdSE", "avgSpan", "converged", "maxMach")}, c.stab, flush=True)
    assert c.stab["events"] == 0
    assert c.cur["U"] == pytest.approx(c.flow["U"])
    assert c.cur["nu"] == pytest.approx(c.flow["nu"], rel=1e-3)  # start-up damping has decayed
    assert math.isfinite(r["cd"]) and 0.2 < r["cd"] < 3.0 and r["maxMach"] < 0.6
    s.destroy()
