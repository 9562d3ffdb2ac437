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


def test_node_harness_freestream_and_cylinder():
    r = subprocess.run([NODE, "windtunnel_vis_amd/node/harness.mjs", "--case=freestream,cylinder", "--valscale=0.5"],
                       cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
