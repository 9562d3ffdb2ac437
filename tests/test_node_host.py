"""Node host (the reference's language): the N-API addon loads under the installed Node and the
ESM host module parses and exposes the reference surface (no GPU calls)."""
import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
pytestmark = pytest.mark.skipif(NODE is None, reason="node not installed")


def _node(code):
    r = subprocess.run([NODE, "-e", code], cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=60)
    return r


def test_addon_exports():
    r = _node("const a=require('./windtunnel_vis_amd/_lib/wt_addon.node'); console.log(JSON.stringify(Object.keys(a)))")
    assert r.returncode == 0, r.stderr
    keys = set(json.loads(r.stdout))
    for k in ["create", "destroy", "setParams", "reset", "clearStats", "uploadFlags", "readFlags", "uploadSdf",
              "setWheels", "step", "stepAndSample", "readPopulations", "uploadPopulations", "voxelizeSdf",
              "voxelizeMesh", "bytesFor", "derive", "exportFields"]:
        assert k in keys


def test_host_module_surface_and_errors():
    code = ("import('./windtunnel_vis_amd/node/solver.mjs').then(m => {"
            " const out = {keys: Object.keys(m), proto: Object.getOwnPropertyNames(m.SolverGPU.prototype),"
            " bytes: m.SolverGPU.bytesFor({nx: 256, ny: 128, nz: 128})};"
            " try { new m.SolverGPU({nx: 0, ny: 0, nz: 0}, {u: 0.05, nu: 0.01}); } catch (e) { out.err = e.message; }"
            " console.log(JSON.stringify(out)); })")
    r = _node(code)
    assert r.returncode == 0, r.stderr
    o = json.loads(r.stdout)
    assert {"SolverGPU", "solverFactory", "D2Q9", "D3Q19"} <= set(o["keys"])
    for k in ["stepAndSample", "encodeSteps", "uploadFlags", "uploadSdf", "setWheels", "readPopulations", "readFlags",
              "setParams", "reset", "clearStats", "destroy", "derive", "exportFields"]:
        assert k in o["proto"]
    assert o["bytes"] == 256 * 128 * 128 * (19 * 8 + 56)
    assert "dims" in o["err"]


def test_js_files_parse_under_installed_node():
    for f in ["solver.mjs", "analysis.mjs", "validation.mjs", "harness.mjs", "export.mjs"]:
        r = subprocess.run([NODE, "--check", os.path.join(ROOT, "windtunnel_vis_amd", "node", f)], stdout=subprocess.PIPE,
                           stderr=subprocess.PIPE, text=True)
        assert r.returncode == 0, (f, r.stderr)


def test_node_writers_match_python_writers(tmp_path):
    """The Node host's csv() / vtk() and the Python host's produce the same bytes."""
    import numpy as np
    from windtunnel_vis_amd import export as E
    from windtunnel_vis_amd.solver import FieldExport
    rng = np.random.default_rng(11)
    dims, stride = (5, 4, 3), 2
    m = 60
    arrs = [rng.normal(0.05, 0.02, (m, 4)).astype(np.float32) for _ in range(3)]
    for a in arrs[:2]:
        a[:, 3] = 1 + rng.normal(0, 1e-3, m)
    f = FieldExport(dims, stride, *arrs)
    np.concatenate([a.reshape(-1) for a in arrs]).astype("<f4").tofile(str(tmp_path / "f.bin"))
    rows = [[1, 0.123456789, "a"], [2.5e-7, float("nan"), "b"], [1234567.5, 6.02214076e23, "c"]]
    js_rows = "[[1,0.123456789,'a'],[2.5e-7,NaN,'b'],[1234567.5,6.02214076e23,'c']]"
    code = ("const fs=require('fs');import('./windtunnel_vis_amd/node/export.mjs').then(m=>{"
            "const b=fs.readFileSync(%r);const a=new Float32Array(b.buffer,b.byteOffset,b.length/4);"
            "const f={dims:[5,4,3],stride:2,vel:a.subarray(0,240),mean:a.subarray(240,480),vort:a.subarray(480,720)};"
            "fs.writeFileSync(%r,m.vtk(f,{dx:0.0125,U:0.06,L:48,rhoRef:1.0003,title:'truck'}));"
            "fs.writeFileSync(%r,m.csv(['t','cd','label'],%s,['run 1']));"
            "console.log(m.strideFor(512*256*256,1e6));})"
            % (str(tmp_path / "f.bin"), str(tmp_path / "f.vtk"), str(tmp_path / "f.csv"), js_rows))
    r = _node(code)
    assert r.returncode == 0, r.stderr
    assert int(r.stdout) == E.stride_for(512 * 256 * 256, 1e6)
    assert (tmp_path / "f.vtk").read_bytes() == E.vtk(f, dx=0.0125, U=0.06, L=48.0, rhoRef=1.0003, title="truck")
    assert (tmp_path / "f.csv").read_text() == E.csv(["t", "cd", "label"], rows, ["run 1"])


def test_node_controller_matches_python_controller():
    """The Node host's RunController and the Python one agree frame by frame on a scripted solver
    (forces as a function of the step count, one unstable sample)."""
    import math
    from windtunnel_vis_amd.controller import RunController
    from windtunnel_vis_amd.solver import SolverStats

    class Fake:
        def __init__(self):
            self.dims = dict(nx=64, ny=32, nz=32)
            self.stepCount = 0

        def setParams(self, p):
            self.last = dict(p)

        def reset(self):
            self.stepCount = 0

        def stepAndSample(self, n):
            self.stepCount += n
            s = self.stepCount
            return SolverStats(0.02 * (1 + 0.1 * math.sin(s / 37.0)), 0.003 * math.cos(s / 53.0), 1e-4, 0.05, 1.0 + 1e-4 * math.sin(s),
                               s == 60, s)

    c = RunController(Fake(), Lcells=40, frontal=90, speed_kmh=180)
    for _ in range(400):
        c.frame(20)
    py = dict(c.readout(), stab=c.stab, cur=c.cur, clock=c.clock, last=c.solver.last)
    code = ("import('./windtunnel_vis_amd/node/controller.mjs').then(async (m) => {"
            " const f = {dims: {nx: 64, ny: 32, nz: 32}, stepCount: 0, setParams(p) { this.last = {...p}; },"
            " reset() { this.stepCount = 0; }, async stepAndSample(n) { this.stepCount += n; const s = this.stepCount;"
            " return {fx: 0.02 * (1 + 0.1 * Math.sin(s / 37)), fy: 0.003 * Math.cos(s / 53), fz: 1e-4, maxU: 0.05,"
            " rhoRef: 1 + 1e-4 * Math.sin(s), unstable: s === 60, step: s}; } };"
            " const c = new m.RunController(f, {Lcells: 40, frontal: 90, speedKmh: 180});"
            " for (let k = 0; k < 400; k++) await c.frame(20);"
            " console.log(JSON.stringify({...c.readout(), stab: c.stab, cur: c.cur, clock: c.clock, last: f.last})); })")
    r = _node(code)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout)
    assert js["stab"] == py["stab"] and py["stab"]["events"] == 1
    for k in ("cd", "cl", "cs", "cdInst", "cdSE", "avgSpan", "cdCorrected", "dragN", "convTime", "maxMach", "blockage"):
        assert js[k] == pytest.approx(py[k], rel=1e-12, abs=1e-15), k
    assert js["converged"] == py["converged"]
    for k in ("U", "nu"):
        assert js["cur"][k] == pytest.approx(py["cur"][k], rel=1e-12)
    for k in ("reSim", "U", "nu", "dx", "dt"):
        assert js["flow"][k] == pytest.approx(py["flow"][k], rel=1e-12)
    assert js["last"]["emaAlpha"] == pytest.approx(py["last"]["emaAlpha"], rel=1e-12)
