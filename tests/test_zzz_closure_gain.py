"""What the trained network gains over plain Smagorinsky on the learned-closure workload (truck +
trailer: fine 512x256x256 reference, coarse 256x128x128 runs), through the whole pipeline of
closure_study.study(): fine run and 2x box-filtered training records, GPU training, output
calibration on two seeds at each of nine scales, evaluation on twelve other seeds for each closure.

Asserted: every run stable, the app's convergence rule (app.ts:447-468) held in every run, and the
learned ensemble's mean C_D closer to the fine C_D than the plain ensemble's by more than two
standard errors of the difference. Reported (the workload's target, not asserted): the learned mean
within 2 % of the fine C_D, and the time to a converged C_D within 2 % against the fine run's time
to its converged C_D.

The module is named to be collected after every other GPU test: it runs for about two minutes.
"""
import json

import pytest

pytestmark = pytest.mark.gpu


def test_learned_closure_gain_truck_trailer():
    from windtunnel_vis_amd import closure_study
    r = closure_study.study(log=lambda m: print(m, flush=True))
    p, l = r["plain"], r["learned"]
    report = dict(cd_fine=r["cd_fine"], plain=[p["mean"], p["se"], p["err"]], learned=[l["mean"], l["se"], l["err"]],
                  gain=r["gain"], gain_se=r["gain_se"], factor=r["output_factor"], within2=l["within2_runs"],
                  within2_median_s=r["learned_within2_median_s"], fine_converged_s=r["fine_converged_s"],
                  speedup=r.get("speedup_vs_fine"))
    print(json.dumps(report), flush=True)
    assert not p["unstable"] and not l["unstable"]
    assert all(s is not None for s in p["converged_step"] + l["converged_step"]), report
    assert r["gain"] > 2 * r["gain_se"], report
