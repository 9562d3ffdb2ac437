"""Writes tests/golden/sedan_c2_oracle.json: C_D and C_L of the sedan LES at 256 x 128 x 128 (the
sedan configuration: u = 0.08, nu = 6e-4, Cs = 0.1, moving belt, rolling wheels, Bouzidi walls, the
app's sponge / inlet-layer / wall defaults) sampled every 250 steps over the first 5000 steps from the
impulsive start, from the FP64 oracle without storage rounding, and their means over the second half
(steps 2500-5000, the reference harness's rule, harness.ts:172-183).

The oracle runs on the GPU voxelizer's geometry (flags, wall distances, wheel table), the inputs the GPU
run gets, so the script has two stages:

  python tests/golden/make_c2_oracle.py export GEOM.npz   # where the library runs (a GPU): geometry
  python tests/golden/make_c2_oracle.py run GEOM.npz      # any machine: the oracle (about 20 minutes
                                                          # on 8 cores)
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
OUT = os.path.join(ROOT, "tests", "golden", "sedan_c2_oracle.json")
DIMS = (256, 128, 128)
U = 0.08
STEPS, CHUNK = 5000, 250


def params(precision="f32"):
    return dict(u=U, nu=6e-4, cs=0.1, ground="moving", sides="freeslip", spongeNu=0.04, spongeStart=0.84, spongeIn=0.03,
                tauWall=0.53, interp=True, collision="regularized", emaAlpha=0.02, precision=precision)


def export(path):
    from windtunnel_vis_amd import SolverGPU
    from windtunnel_vis_amd import geometry as G
    s = SolverGPU(DIMS, params())
    body = G.sedan(DIMS)
    s.setWheels(body.wheels)
    info = s.voxelizeSdf(body.prims())
    wheels = [{k: ([float(x) for x in v] if np.ndim(v) else float(v)) for k, v in w.items()} for w in s.wheels_]
    np.savez_compressed(path, flags=s.readFlags(), sdf=s.readSdf(), frontal=info["frontal"], wheels=json.dumps(wheels))
    s.destroy()


def series(solver, frontal):
    """C_D, C_L every CHUNK steps over STEPS steps (A = frontal cells, rho = 1)."""
    from windtunnel_vis_amd.analysis import coefficients
    cd, cl = [], []
    for _ in range(STEPS // CHUNK):
        r = solver.stepAndSample(CHUNK)
        c = coefficients(r.fx, r.fy, r.fz, U, frontal)
        cd.append(c["cd"])
        cl.append(c["cl"])
    return cd, cl


def second_half(v):
    return float(np.mean(v[len(v) // 2:]))


def run(path):
    from oracle.oracle import OracleSolver
    g = np.load(path)
    o = OracleSolver(DIMS, params())
    o.setWheels(json.loads(str(g["wheels"])))
    o.uploadFlags(g["flags"])
    o.uploadSdf(g["sdf"])
    t0 = time.time()
    cd, cl = series(o, int(g["frontal"]))
    o.destroy()
    out = dict(dims=list(DIMS), steps=STEPS, every=CHUNK, frontal=int(g["frontal"]), cd=cd, cl=cl,
               cd_second_half=second_half(cd), cl_second_half=second_half(cl), oracle="FP64 CPU restatement, no storage rounding",
               seconds=round(time.time() - t0, 1))
    json.dump(out, open(OUT, "w"), indent=1)
    print(json.dumps({k: out[k] for k in ("cd_second_half", "cl_second_half", "seconds")}))


if __name__ == "__main__":
    {"export": export, "run": run}[sys.argv[1]](sys.argv[2])
