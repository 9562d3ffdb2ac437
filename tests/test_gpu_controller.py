"""The run controller drives the MI355X solver: a sphere in a 96x48x48 tunnel, 120 frames of 20
steps with the start-up ramps, EMA schedule and watchdog active."""
import math

import pytest

pytestmark = pytest.mark.gpu


def test_controller_driven_run_matches_oracle():
    """The run controller on the GPU solver and on the FP64 oracle, frame by frame: every frame sets
    u, nu (start-up ramp and damping), Cs, the floor, the EMA weight and interp through setParams
    (app.ts:792-810 -> SolverGPU.ts:155-187; the GPU rebuilds its pull table and wall links each
    time), and mid-run the interpolated walls are switched off and Cs raised. FP32 storage: the
    statistics of every frame and the final populations agree with the oracle's (1e-5 relative)."""
    from oracle.oracle import OracleSolver
    from tests.cases import box_body_3d, make_case
    from windtunnel_vis_amd import SolverGPU
    from windtunnel_vis_amd.controller import RunController
    import numpy as np
    dims = (32, 16, 12)
    flags, sdf, wheels = box_body_3d(dims)
    _, params, _, _, _ = make_case("d3_full")
    p = dict(params, precision="f32")
    runs = []
    for make in (lambda: SolverGPU(dims, p), lambda: OracleSolver(dims, p)):
        s = make()
        s.setWheels(wheels)
        s.uploadFlags(flags)
        s.uploadSdf(sdf)
        c = RunController(s, Lcells=8, frontal=20, speed_kmh=100, reOverride=200, ground="moving")
        hist = []
        for f in range(60):
            if f == 30:
                c.interp = False
                c.flowChanged()
            if f == 45:
                c.cs = 0.15
                c.flowChanged()
            st = c.frame(5)
            hist.append((st.step, st.fx, st.fy, st.maxU, st.rhoRef))
        runs.append((s, c, np.array(hist)))
    (g, cg, hg), (o, co, ho) = runs
    try:
        assert np.array_equal(hg[:, 0], ho[:, 0])
        scale = max(1e-3, float(np.abs(ho[:, 1:3]).max()))
        err_f = float(np.abs(hg[:, 1:3] - ho[:, 1:3]).max()) / scale
        err_u = float(np.abs(hg[:, 3:] - ho[:, 3:]).max())
        fg, fo = g.readPopulations().astype(np.float64), o.readPopulations()
        err_p = float((np.abs(fg - fo) / np.abs(fo)).max())
        print(f"controller-driven run: force {err_f:.2e}, maxU/rhoRef {err_u:.2e}, populations {err_p:.2e}; "
              f"cd gpu {cg.readout()['cd']:.6f} oracle {co.readout()['cd']:.6f}", flush=True)
        assert err_f < 1e-4 and err_u < 1e-6 and err_p < 1e-5
        assert cg.stab == co.stab
        # the controllers' readouts (averaging windows, EMA, convergence) follow the same samples
        rg, ro = cg.readout(), co.readout()
        for k in ("cd", "cl", "cdSE", "converged"):
            a, b = rg[k], ro[k]
            if isinstance(a, float) and math.isnan(a):
                assert math.isnan(b), k
            else:
                assert a == pytest.approx(b, rel=1e-4, abs=1e-6), k
    finally:
        g.destroy()
        o.destroy()


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


def test_fences_follow_the_queue():
    """wt_fence / wt_fence_done / wt_fence_wait: a fence passes once the work queued before it is
    done (waited for, or after a sync); tickets count up; the ring of 16 records is reused without
    losing older tickets; a ticket not yet issued is refused. setParams with only a speed change
    leaves the queue running (the fence after it still follows the queued steps)."""
    from windtunnel_vis_amd import SolverGPU
    s = SolverGPU((64, 32, 24), dict(u=0.05, nu=0.02, cs=0.1, ground="moving", precision="f16"))
    try:
        s.step(50)
        s.setParams(dict(u=0.06))
        t1 = s.fence()
        s.fenceWait(t1)
        assert s.fenceDone(t1)
        ts = []
        for k in range(40):
            s.step(2)
            ts.append(s.fence())
        assert ts == list(range(t1 + 1, t1 + 41))
        s.fenceWait(ts[-1])
        assert all(s.fenceDone(t) for t in ts)  # older tickets, their records reused, are passed too
        s.step(10)
        t2 = s.fence()
        s.sync()
        assert s.fenceDone(t2)
        with pytest.raises(RuntimeError):
            s.fenceDone(t2 + 1)
    finally:
        s.destroy()
