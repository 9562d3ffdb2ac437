"""The Node host drives the GPU path end to end: its results match the Python host bit for bit
(same C ABI), and the reference's validation harness cases pass through it."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
pytestmark = [pytest.mark.gpu, pytest.mark.skipif(NODE is None, reason="node not installed")]


def test_node_matches_python_host():
    from windtunnel_vis_amd import SolverGPU
    r = subprocess.run([NODE, "tests/node/parity_check.mjs", "f16"], cwd=ROOT, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    o = json.loads(r.stdout)
    dims = (24, 12, 10)
    s = SolverGPU(dims, dict(u=0.06, nu=0.01, cs=0.1, ground="moving", sides="freeslip", spongeNu=0.04, spongeStart=0.8,
                             spongeIn=0.15, tauWall=0.53, emaAlpha=0.1, precision="f16"))
    flags = np.zeros((10, 12, 24), np.uint32)
    flags[3:7, 4:8, 9:13] = 1
    s.uploadFlags(flags.reshape(-1))
    st = s.stepAndSample(100)
    f = s.readPopulations()
    assert o["stepCount"] == 100 and o["precision"] == "f16"
    assert o["stats"]["fx"] == st.fx and o["stats"]["fy"] == st.fy
    assert np.array_equal(np.array(o["first"], dtype=np.float32), f.reshape(-1)[:8])
    assert abs(o["sum"] - float(f.astype(np.float64).sum())) < 1e-3
    s.derive()
    e = s.exportFields(2)
    assert tuple(o["ex"]["dims"]) == e.dims
    np.testing.assert_array_equal(np.array(o["ex"]["vort8"], dtype=np.float32), e.vort.reshape(-1)[:8])
    for k, a in (("vel", e.vel), ("mean", e.mean), ("vort", e.vort)):
        assert abs(o["ex"][k] - float(a.astype(np.float64).sum())) <= 1e-9 * max(1.0, float(np.abs(a).sum()))
    # tracers through the Node host equal the Python host's (the same kernels on the same state)
    from windtunnel_vis_amd import Tracers
    t = o["tracers"]
    vel, mean = s.probe(np.array([[2.5, 6.0, 5.0], [20.25, 3.5, 7.75], [11.0, 0.2, 9.9]]))
    np.testing.assert_array_equal(np.array(t["vel"], np.float32), vel.reshape(-1))
    np.testing.assert_array_equal(np.array(t["mean"], np.float32), mean.reshape(-1))
    sl = s.streamlines(dict(a=(1.5, 2.0, 3.0), b=(1.5, 10.0, 3.0), d=(0.0, 0.0, 2.0)), 6, 12, 0.5, perLine=3)
    np.testing.assert_array_equal(np.array(t["sl"], np.float32), sl.reshape(-1))
    assert (sl[:, 0, 3] > 0).all()  # seeds start in the flow
    tr = Tracers(s, 64, trail=4, emitter="inlet", maxAge=50)
    tr.update(0, advance=False, fill=True, seed=1)
    tr.update(2, seed=2)
    tr.update(2, seed=3)
    state, hist, head = tr.read()
    tr.destroy()
    np.testing.assert_array_equal(np.array(t["state"], np.float32), state.reshape(-1))
    assert t["head"] == head and abs(t["trail"] - float(hist.astype(np.float64).sum())) <= 1e-9 * max(1.0, float(np.abs(hist).sum()))


def test_node_async_frames_and_calls_during_a_sample():
    """encodePost / flushReads (the reference's out-of-order-tolerant frame reads): every delivered
    read carries its post's step, in order, and the last equals a synchronous sample of the same
    state; a read posted after a flush and completed before the next one goes to the next one's
    callback; the reference's throttled frame loop (at most two frames in flight, counted with
    onSubmittedWorkDone, setParams every frame) keeps its bound and delivers its reads. encodeSteps
    during an un-awaited stepAndSample runs after it (the handle lock): the results
    equal a strictly sequential run on a second solver. destroy() during a sample is deferred: the
    sample still resolves and later calls throw."""
    r = subprocess.run([NODE, "tests/node/async_frames.mjs"], cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    o = json.loads(r.stdout)
    print(o)
    assert o["burst"][:4] == [10, 20, 30, 40] and o["burst"] == sorted(o["burst"])
    assert set(o["burst"]) <= {10, 20, 30, 40, 50, 60}
    steps = [f["step"] for f in o["frames"]]
    assert steps == o["burst"] + [60]
    assert o["latest"] == 60 and o["syncStep"] == 60
    assert o["frames"][-1]["fx"] == o["syncFx"]
    assert o["late"] == [["first", 10], ["second", 20]]
    th = o["throttled"]
    assert th["frames"] == 40 and th["step"] == 200 and 1 <= th["maxInFlight"] <= 2
    assert th["reads"] == sorted(th["reads"]) and len(th["reads"]) >= 36 and th["reads"][-1] == 200
    race = o["race"]
    assert race["a1"] == race["b1"] and race["a2"] == race["b2"] and race["a2"][0] == 230
    assert o["destroyed"] == {"step": 300, "unstable": False, "threwAfterDestroy": True}


def test_python_host_async_reads_match_sync_sample():
    from windtunnel_vis_amd import SolverGPU
    import time
    dims = (24, 12, 10)
    s = SolverGPU(dims, dict(u=0.06, nu=0.01, cs=0.1, ground="moving", precision="f16"))
    got = []

    def drain():
        t0 = time.time()
        while True:
            done, pending = s.flushReads()
            got.extend(done)
            if pending == 0 or time.time() - t0 > 10:
                return pending

    queued = []
    for k in range(6):
        s.step(10)
        queued.append(s.encodePost(True))
        got += s.flushReads()[0]
    assert drain() == 0
    assert queued[:4] == [True] * 4
    steps = [g.step for g in got]
    assert steps == [10 * (k + 1) for k in range(6) if queued[k]]
    assert s.encodePost(True)
    assert drain() == 0
    assert got[-1].step == 60
    st = s.stepAndSample(0)
    assert (got[-1].fx, got[-1].fy, got[-1].maxU, got[-1].rhoRef) == (st.fx, st.fy, st.maxU, st.rhoRef)
    assert s.latest.step == 60
    s.destroy()


def test_node_harness_freestream_and_cylinder():
    r = subprocess.run([NODE, "windtunnel_vis_amd/node/harness.mjs", "--case=freestream,cylinder", "--valscale=0.5"],
                       cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
