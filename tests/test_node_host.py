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
              "voxelizeMesh", "bytesFor"]:
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
              "setParams", "reset", "clearStats", "destroy"]:
        assert k in o["proto"]
    assert o["bytes"] == 256 * 128 * 128 * (19 * 8 + 56)
    assert "dims" in o["err"]


def test_js_files_parse_under_installed_node():
    for f in ["solver.mjs", "analysis.mjs", "validation.mjs", "harness.mjs"]:
        r = subprocess.run([NODE, "--check", os.path.join(ROOT, "windtunnel_vis_amd", "node", f)], stdout=subprocess.PIPE,
                           stderr=subprocess.PIPE, text=True)
        assert r.returncode == 0, (f, r.stderr)
