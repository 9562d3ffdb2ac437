"""Writes tests/golden/validation_oracle.json: the reference's validation cases (validation.ts) run on
the FP64 CPU oracle (the restatement of the reference's kernels), at the reduced size the reference's
CI runs them (valscale 0.5, scripts/harness.mjs:11-15):

* cylinder (D2Q9, D = 12, Re = 100, Bouzidi): St, mean C_D, C_L amplitude, cycles;
* Ahmed 25 degrees (D3Q19 LES, L = 32 cells): mean C_D and its standard error (the preset mesh
  through the mesh voxelizer's restatement).

The GPU tests compare the MI355X solver's values with these (the error against the FP64 oracle for
St and C_D). Run from the repository root:
  python tests/golden/make_validation_oracle.py [--scale=0.5] [cases...]
Valscale 0.5 takes about five minutes on 8 cores; the Ahmed case at valscale 1 (L = 64, the GPU
test's size) about an hour and a half on 5.
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import OracleSolver  # noqa: E402
from windtunnel_vis_amd import validation as V  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "validation_oracle.json")


def main(cases, scale=0.5):
    make = lambda d, p: OracleSolver(d, p)  # noqa: E731
    for name in cases:
        t0 = time.time()
        r = V.CASES[name](make, scale=scale)
        m = {k: (float(v) if isinstance(v, float) else v) for k, v in r["metrics"].items()}
        key = "%s_valscale_%g" % (name, scale)
        entry = dict(metrics=m, passed=bool(r["passed"]), oracle="FP64 CPU restatement", seconds=round(time.time() - t0, 1))
        print(key, entry, flush=True)
        data = json.load(open(OUT)) if os.path.exists(OUT) else {}
        data[key] = entry
        json.dump(data, open(OUT, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    args = sys.argv[1:]
    scale = 0.5
    if args and args[0].startswith("--scale="):
        scale = float(args.pop(0).split("=", 1)[1])
    main(args or ["cylinder", "ahmed"], scale)
