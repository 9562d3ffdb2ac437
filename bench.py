"""Benchmark: D3Q19 Smagorinsky LES MLUPS and HBM-roofline fraction on one MI355X.

Headline workload (default): the plain D3Q19 LES at 512x256x256 with FP16 deviation storage,
Bouzidi walls, moving belt at u = 0.08 with rolling wheels, Cs = 0.1, around the procedural sedan
(rounded box 85x22x32 cells, corner radius 6, four wheels r = 6, width 5, ride height 3, nose at
0.34 nx). At this size the 2.5 GB ping-pong is ten times the 256 MB memory-side
cache, so the HBM roofline fraction is the memory system's, not the cache's. One step = one fused
pull-stream + collide pass over the whole grid, one launch of the step kernel: some of its blocks
update the cells next to the body from precomputed links, the rest stream the bulk.

Prints ONE JSON line. `value` is MLUPS over the timed steps with the populations resident in HBM.
`roofline` prices the step kernel at 2*Q*s algorithmic bytes per fluid cell, divided by its average
launch time measured with HIP events on the solver's stream, against the 8 TB/s HBM peak; `traffic`
is the HBM bytes per launch from the committed rocprofv3 counter summary of the same workload.
`cpu_baseline` times the CPU restatement on a bounded sample of the same workload. The 256x128x128
line (the sedan LES configuration with its 5000-step C_D / C_L) and the other workloads are extra keys.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "D3Q19 LES MLUPS + % HBM roofline; learned vs plain time-to-C_D at 2% error"
HBM_PEAK_GBS = 8000.0


def workload_params(u=0.08, precision="f16"):
    # config: Cs 0.1, u 0.08, moving belt, interp; nu = 6e-4 (the app's 3D floor) and the app's
    # sponge / wall defaults (app.ts:227-231)
    return dict(u=u, nu=6e-4, cs=0.1, ground="moving", sides="freeslip", spongeNu=0.04, spongeStart=0.84,
                spongeIn=0.03, tauWall=0.53, interp=True, collision="regularized", emaAlpha=0.02, precision=precision)


def make_sedan_solver(dims, precision="f16"):
    from windtunnel_vis_amd import SolverGPU
    from windtunnel_vis_amd import geometry as G
    s = SolverGPU(dims, workload_params(precision=precision))
    body = G.sedan(dims)
    s.setWheels(body.wheels)
    info = s.voxelizeSdf(body.prims())
    return s, body, info


def traffic_from_profiles(name):
    """HBM bytes per bulk launch from a committed PMC summary (FETCH_SIZE + WRITE_SIZE, scaled)."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def run_gpu(dims, steps, warmup, precision="f16", chunk=None):
    """Time `steps` steps. chunk None: launch configuration from the autotuner; else that item-run
    length (4 = plain order, 0 = one contiguous eighth per XCD)."""
    from windtunnel_vis_amd import tuning
    s, body, info = make_sedan_solver(dims, precision)
    tune = None
    if chunk is None:
        tune = tuning.autotune(s, steps=20)
    else:
        s.setTuning(0, chunk)
    q = s.lattice.q
    sbytes = 2 if s.precision != "f32" else 4
    n = s.n
    s.step(warmup)
    s.sync()
    # events around the timed launches only: events bound to each launch would add queue time
    # between the launches (a few microseconds each) to the very region being timed
    s.profileEnable("span")
    s.sync()
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    t1 = time.perf_counter()
    k_ms, launches, _ = s.profileRead()
    s.profileEnable(False)
    # per-launch event durations over a short second stretch, as a cross-check of the span average
    s.profileEnable("launch")
    s.step(min(steps, 200))
    k_ms_launch, launches_launch, _ = s.profileRead()
    s.profileEnable(False)
    st = s.stepAndSample(0)
    bulk_cells, wall_cells = s.workSplit()
    wall = t1 - t0
    mlups = n * steps / wall / 1e6
    k_avg_s = k_ms / 1e3 / max(launches, 1)
    fluid = n - info["solidCells"]
    k_bytes = fluid * 2 * q * sbytes
    achieved = k_bytes / k_avg_s / 1e9
    return dict(solver=s, body=body, info=info, mlups=mlups, wall=wall, k_avg_ms=k_avg_s * 1e3, fluid=fluid,
                k_avg_ms_launch_events=k_ms_launch / max(launches_launch, 1),
                bulk_cells=bulk_cells, wall_cells=wall_cells, k_bytes=k_bytes, achieved_gbs=achieved,
                unstable=st.unstable, maxU=st.maxU, q=q, sbytes=sbytes, tune=tune)


def shard_bench(steps, warmup, precision="f16", parts=4):
    """One GPU's share of the 512x256x256 sedan domain cut into `parts` x-slabs: the slab that holds
    the body (x in [128, 256) for four parts) as its own 128x256x256 grid, with the cut planes as the
    domain's inlet / outlet faces (the halo exchange that would replace them is not built). Measures
    the step kernel at the shard's real shape (rows of 128 cells instead of 512)."""
    from windtunnel_vis_amd import SolverGPU, tuning
    from windtunnel_vis_amd import geometry as G
    full = (512, 256, 256)
    w = full[0] // parts
    x0 = w * int(0.34 * full[0] // w)  # the slab holding the body's centre
    dims = (w, full[1], full[2])
    s = SolverGPU(dims, workload_params(precision=precision))
    body = G.sedan(full)
    body.translate((-x0, 0, 0))
    s.setWheels(body.wheels)
    info = s.voxelizeSdf(body.prims())
    tune = tuning.autotune(s, steps=20)
    s.step(warmup)
    s.profileEnable("span")
    s.sync()
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    wall = time.perf_counter() - t0
    k_ms, launches, _ = s.profileRead()
    s.profileEnable(False)
    st = s.stepAndSample(0)
    s.destroy()
    k_avg = k_ms / max(launches, 1)
    fluid = s.n - info["solidCells"]
    sb = 4 if precision == "f32" else 2
    return {"slab": "x in [%d, %d) of 512x256x256 (%d x-slabs)" % (x0, x0 + w, parts), "grid": list(dims),
            "solid_cells": info["solidCells"], "mlups": round(s.n * steps / wall / 1e6, 1), "avg_launch_ms": round(k_avg, 5),
            "kernel_frac": round(fluid * 2 * 19 * sb / (k_avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "chunk": tune["chunk"],
            "unstable": bool(st.unstable)}


def learned_step_bench(dims, steps, warmup, plain_ms, precision="f16"):
    """The step with the learned closure (k_step_mlp: BF16 MFMA inference fused in the step) on the
    same sedan domain and launch configuration as the plain line (2 cells per lane, plain item
    order). The network is an untrained one of the trained network's shape and normalisation: the
    cost does not depend on the weights. Timed as run_gpu times the plain step."""
    from windtunnel_vis_amd import closure
    s, body, info = make_sedan_solver(dims, precision)
    s.setTuning(0, 4)
    net = closure.Network.initial(nu_scale=1e-4)
    net.sigma = np.full(8, 3.0, np.float32)
    closure.attach(s, net)
    s.step(warmup)
    s.profileEnable("span")
    s.sync()
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    wall = time.perf_counter() - t0
    k_ms, launches, _ = s.profileRead()
    s.profileEnable(False)
    st = s.stepAndSample(0)
    s.destroy()
    k_avg = k_ms / max(launches, 1)
    fluid = s.n - info["solidCells"]
    sb = 4 if precision == "f32" else 2
    return {"kernel": "k_step_mlp<SF16,2,regularized>", "mlups": round(s.n * steps / wall / 1e6, 1),
            "avg_launch_ms": round(k_avg, 5), "hbm_frac": round(fluid * 2 * 19 * sb / (k_avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "time_vs_plain_step": round(k_avg / plain_ms, 3) if plain_ms else None, "steps": steps,
            "unstable": bool(st.unstable)}


def precision_stress(dims=(1024, 512, 512), steps=20, warmup=5):
    """The sports-wedge setup (300x90x45 cells, corner radius 10, u = 0.1) at 1024x512x512 in each
    storage format: MLUPS and the step kernel's HBM roofline fraction (autotuned launch)."""
    from windtunnel_vis_amd import SolverGPU, tuning
    from windtunnel_vis_amd import geometry as G
    out = {"grid": list(dims), "body": "sports wedge 300x90x45, corner 10", "u": 0.1}
    for prec in ("f32", "f16", "bf16"):
        s = SolverGPU(dims, workload_params(u=0.1, precision=prec))
        body = G.sports_wedge(dims)
        info = s.voxelizeSdf(body.prims())
        tune = tuning.autotune(s, steps=10)
        s.step(warmup)
        s.profileEnable("span")
        s.sync()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        wall = time.perf_counter() - t0
        k_ms, launches, _ = s.profileRead()
        st = s.stepAndSample(0)
        sb = 4 if prec == "f32" else 2
        fluid = s.n - info["solidCells"]
        k_avg = k_ms / 1e3 / max(launches, 1)
        mlups = s.n * steps / wall / 1e6
        out[prec] = {"mlups": round(mlups, 1), "ms_per_step": round(wall * 1e3 / steps, 4),
                     "avg_launch_ms": round(k_avg * 1e3, 4),
                     "kernel_frac": round(fluid * 2 * 19 * sb / k_avg / 1e9 / HBM_PEAK_GBS, 4),
                     "step_frac": round(mlups * 1e6 * 2 * 19 * sb / (HBM_PEAK_GBS * 1e9), 4),
                     "chunk": tune["chunk"], "cells_per_thread": tune["cellsPerThread"], "unstable": bool(st.unstable)}
        s.destroy()
    return out


def copy_probe(nbytes=2 << 30, iters=5):
    from windtunnel_vis_amd import _native as N
    lib = N.load()
    g = N.C.c_double()
    N.check(lib.wt_copy_probe(int(nbytes), int(iters), N.C.byref(g)))
    return g.value


def physics(s, body, info, u, total=5000, max_steps=40000, chunk=250):
    """C_D / C_L of a run from the init state, read two ways: as the configuration states it, the mean
    over the second half of a `total`-step run (harness.ts:172-183), and as the app reads them
    (app.ts:447-468): samples from one convective time on (the app's averaging start after a
    reset), the window of the last 4 L/U, the run continued past `total` until the convergence rule
    holds there (span >= 3 L/U, SE < max(2 % |mean|, 0.005)) or `max_steps`."""
    from windtunnel_vis_amd.analysis import coefficients, converged, windowStats
    s.reset()
    t, cd, cl = [], [], []
    half_d, half_l = [], []
    ok, unstable, step = False, False, 0
    L = body.length
    while step < max_steps and not (ok and step >= total):
        r = s.stepAndSample(chunk)
        step += chunk
        if r.unstable:
            unstable = True
            break
        c = coefficients(r.fx, r.fy, r.fz, u, info["frontal"])
        if total / 2 < step <= total:
            half_d.append(c["cd"])
            half_l.append(c["cl"])
        if step < L / u:
            continue
        t.append(step)
        cd.append(c["cd"])
        cl.append(c["cl"])
        ok, w = converged(t, cd, L, u)
    ok, w = converged(t, cd, L, u)
    wl = windowStats(t, cl, 4 * L / u)
    nan = float("nan")
    return dict(cd=w["mean"], cd_se=w["se"], cl=wl["mean"], converged=ok, steps=step, unstable=unstable,
                cd_half=float(np.mean(half_d)) if half_d else nan, cl_half=float(np.mean(half_l)) if half_l else nan,
                total=total, frontal=info["frontal"], L=L)


def sweep_bench(steps=500, max_steps=16000, chunk=250):
    """Batched yaw / ride-height sweep: 24 uneven D3Q19 FP16 domains (autotuned launch configuration
    per domain) stepped by one launch per step in 100-step hipGraph chunks. Aggregate MLUPS (batched
    with graphs, batched direct launches, and the same domains stepped one after another), then
    per-case C_D / C_L read as the app reads them (samples from one convective time on, window of the
    last 4 L/U), the batch stepped until the convergence rule holds in every case or `max_steps`."""
    from windtunnel_vis_amd import BatchGPU, sweep, tuning
    from windtunnel_vis_amd.analysis import coefficients
    cases = sweep.sweep_cases()
    made = sweep.make_solvers(cases)
    solvers = [m[0] for m in made]
    for s_ in solvers:  # launch configuration per domain from the autotuner
        tuning.autotune(s_, steps=10)
    cells = sum(s.n for s in solvers)
    batch = BatchGPU(solvers, chunk=100, graph=True)
    batch.run(100)
    batch.sync()

    def timed(fn):
        t0 = time.perf_counter()
        fn()
        return time.perf_counter() - t0

    def run_batch():
        batch.run(steps)
        batch.sync()

    t_graph = timed(run_batch)
    batch.setChunk(100, graph=False)
    t_direct = timed(run_batch)
    batch.setChunk(100, graph=True)

    def run_serial():
        for s_ in solvers:
            s_.step(steps)
        for s_ in solvers:
            s_.sync()

    batch.destroy()
    t_serial = timed(run_serial)
    batch = BatchGPU(solvers, chunk=100, graph=True)
    for s_ in solvers:
        s_.reset()
    u = sweep.sweep_params()["u"]
    from windtunnel_vis_amd.analysis import converged, windowStats
    series = [([], [], []) for _ in solvers]
    unstable = [False] * len(solvers)
    done = [False] * len(solvers)
    conv_step = [None] * len(solvers)
    step = 0
    while step < max_steps and not all(done[i] or unstable[i] for i in range(len(solvers))):
        batch.run(chunk)
        step += chunk
        for i, (r, (s_, body, info)) in enumerate(zip(batch.stats(), made)):
            unstable[i] |= r.unstable
            if step < body.length / u:
                continue
            c = coefficients(r.fx, r.fy, r.fz, u, info["frontal"])
            t_, d_, l_ = series[i]
            t_.append(step)
            d_.append(c["cd"])
            l_.append(c["cl"])
            ok, _ = converged(t_, d_, body.length, u)
            if ok and not done[i]:
                done[i], conv_step[i] = True, step
    batch.destroy()
    per_case = []
    for c, (s_, body, info), (t_, d_, l_), u_, cs_ in zip(cases, made, series, unstable, conv_step):
        ok, w = converged(t_, d_, body.length, u)
        wl = windowStats(t_, l_, 4 * body.length / u)
        rd = lambda v: round(v, 4) if v == v else None  # noqa: E731
        per_case.append({"dims": list(c["dims"]), "body": c["kind"], "yaw": c["yaw"], "ride": c["ride"],
                         "cd": rd(w["mean"]), "cd_se": rd(w["se"]), "cl": rd(wl["mean"]), "converged_step": cs_,
                         "converged_at_end": ok, "unstable": u_})
    for s_ in solvers:
        s_.destroy()
    return {"domains": len(cases), "cells": cells, "steps": steps,
            "mlups_batched_graph": round(cells * steps / t_graph / 1e6, 1),
            "mlups_batched_direct": round(cells * steps / t_direct / 1e6, 1),
            "mlups_serial_per_domain": round(cells * steps / t_serial / 1e6, 1),
            "cd_steps": step, "cases_converged": sum(c["converged_step"] is not None for c in per_case), "cases": per_case}


def closure_bench():
    """Learned vs plain time-to-C_D: the truck + trailer at 512x256x256 (plain Smagorinsky, the
    reference C_D and the training data) and at 256x128x128 with plain Smagorinsky and with the
    trained network (windtunnel_vis_amd.closure_study)."""
    from windtunnel_vis_amd import closure_study
    r = closure_study.study()
    rd = lambda v: None if v is None or v != v else round(v, 4)  # noqa: E731

    def ens(x):
        return {"cd": [rd(v) for v in x["cd"]], "mean": rd(x["mean"]), "se": rd(x["se"]), "err_vs_fine": rd(x["err"]),
                "converged_at_end": x["converged"], "converged_s": [rd(v) for v in x["converged_wall"]],
                "within_2pct_of_fine_s": [rd(v) for v in x["within2_wall"]], "within_2pct_runs": x["within2_runs"],
                "mlups": round(x["mlups"], 1), "unstable": x["unstable"]}

    f = r["fine"]
    return {"body": "truck + trailer", "fine": {"cd": rd(r["cd_fine"]), "cd_window": rd(f["cd_mean"]), "converged": f["converged"],
                                                "converged_s": rd(f["converged_wall"]), "mlups": round(f["mlups"], 1)},
            "coarse_plain": ens(r["plain"]), "coarse_learned": ens(r["learned"]),
            "gain": rd(r["gain"]), "gain_se": rd(r["gain_se"]), "output_factor": rd(r["output_factor"]),
            "calibration": [{"factor": c["scale"], "cd": [rd(v) for v in c["cd"]], "cd_mean": rd(c["mean"])} for c in r["calibration"]],
            "learned_within_2pct_median_s": rd(r["learned_within2_median_s"]),
            "speedup_vs_fine": rd(r.get("speedup_vs_fine")), "training_records": r["records"],
            "train_seconds": round(r["train_seconds"], 2), "train_loss": [round(x, 5) for x in r["train_loss"]],
            "smag_floor": r["network"].smag_floor, "nu_scale": r["network"].nu_scale,
            "note": "coarse runs of %d steps from seeded perturbed starts (1 %% of U), C_D over each run's second half; "
                    "calibration seeds %s, evaluation seeds %s; gain = |plain mean - fine| - |learned mean - fine|"
                    % (r["coarse_steps"], r["calib_seeds"], r["eval_seeds"])}


def cpu_threads():
    """Threads for the CPU baseline: every CPU this process may run on, capped by OMP_NUM_THREADS
    when that is set (a host that gives each GPU's job a share of its cores sets it to that share)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    cap = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(avail, int(cap))) if cap and cap.isdigit() else avail


def cpu_baseline(solver, sample_steps, threads):
    """The CPU restatement (FP32 arithmetic path with FP16 store rounding, OpenMP) on the same domain:
    the GPU solver's own flags (wheel tags included), wall distances (Bouzidi links) and wheels,
    timed over a few steps."""
    from oracle.oracle import OracleSolver
    dims = (solver.dims["nx"], solver.dims["ny"], solver.dims["nz"])
    p = dict(solver.params_)
    p["precision"] = "f32"
    o = OracleSolver(dims, p, storage="f16", fp32=True, threads=threads)
    o.setWheels(solver.wheels_)
    o.uploadFlags(solver.readFlags())
    o.uploadSdf(solver.readSdf())
    o.step(1)  # init + first touch
    t0 = time.perf_counter()
    o.step(sample_steps)
    dt = time.perf_counter() - t0
    o.destroy()
    n = dims[0] * dims[1] * dims[2]
    return dict(value=n * sample_steps / dt / 1e6, seconds=dt)


def roofline_block(r, copy_gbs, dims, traffic_name):
    return {"bound": "hbm", "achieved": round(r["achieved_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(r["achieved_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles(traffic_name),
            "kernel": "k_step<D3Q19,SF16,2,regularized>", "bytes_per_launch": r["k_bytes"],
            "avg_launch_ms": round(r["k_avg_ms"], 5), "fluid_cells_per_launch": r["fluid"],
            "timing": "HIP event pair on the solver's stream around the timed launches: busy span / launches",
            "avg_launch_ms_events_per_launch": round(r["k_avg_ms_launch_events"], 5),
            "measured_copy_gbs": round(copy_gbs, 1)}


_T0 = time.perf_counter()


def progress(what):
    """One progress line per finished part on stderr (stdout carries only the result line)."""
    print("[bench %6.1f s] %s" % (time.perf_counter() - _T0, what), file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--dims", default="512,256,256", help="headline grid (the sedan LES workload)")
    ap.add_argument("--precision", default="f16")
    ap.add_argument("--no-physics", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=0, help="CPU-baseline sample steps (0 = auto)")
    ap.add_argument("--no-small", action="store_true", help="skip the 256x128x128 line (its physics and the learned step)")
    ap.add_argument("--no-large", action="store_true", help="skip the x-slab shard line")
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched 24-domain sweep")
    ap.add_argument("--no-closure", action="store_true", help="skip the learned-closure study")
    ap.add_argument("--no-stress", action="store_true", help="skip the 1024x512x512 precision stress")
    ap.add_argument("--profile-run", action="store_true",
                    help="short run for counter collection: copy probe + the headline workload only")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="bytes of the copy probe in MiB (0: 256 MiB with --profile-run, else 2048)")
    ap.add_argument("--no-learned-step", action="store_true", help="skip timing the learned-closure step")
    ap.add_argument("--learned-step", action="store_true", help="time the learned-closure step also with --profile-run")
    ap.add_argument("--chunk", type=int, default=None,
                    help="fixed item-run length instead of the autotuner (4 = plain order, 0 = eighths)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    dims = tuple(int(v) for v in args.dims.split(","))
    small = (256, 128, 128)
    if args.profile_run:
        args.no_physics = args.no_cpu = args.no_sweep = args.no_closure = args.no_stress = args.no_large = True
        args.no_small = not args.learned_step
    copy_mib = args.copy_mib or (256 if args.profile_run else 2048)
    copy_gbs = copy_probe(copy_mib << 20, 3 if args.profile_run else 5)
    progress("copy probe %.0f GB/s" % copy_gbs)
    r = run_gpu(dims, args.steps, args.warmup, args.precision, chunk=args.chunk)
    progress("%dx%dx%d step: %.1f MLUPS" % (dims + (r["mlups"],)))
    out = {
        "metric": METRIC, "value": round(r["mlups"], 1), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(r["wall"] * 1e3 / args.steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "D3Q19 Smagorinsky LES %dx%dx%d, %s deviation storage, procedural sedan SDF, "
                               "Bouzidi walls, moving belt + rolling wheels, u=0.08, Cs=0.1" % (dims + (args.precision.upper(),)),
                   "grid": list(dims), "storage": args.precision, "collision": "regularized", "closure": "smagorinsky"},
    }
    out["roofline"] = roofline_block(r, copy_gbs, dims, "pmc_step_d3q19_%s_%dx%dx%d.txt" % ((args.precision,) + dims))
    out["step_roofline_frac"] = round(r["mlups"] * 1e6 * 2 * r["q"] * r["sbytes"] / (HBM_PEAK_GBS * 1e9), 4)
    out["work_split"] = {"bulk_fluid_cells": r["bulk_cells"], "wall_cells": r["wall_cells"]}
    if r["tune"]:
        t = r["tune"]
        out["tuning"] = {"chunk": t["chunk"], "cells_per_thread": t["cellsPerThread"], "cached": t["cached"],
                         "trials_ms_per_step": {"%d/%d" % (e["cellsPerThread"], e["chunk"]): round(e["msPerStep"], 5)
                                                for e in t["trials"]}}
    out["unstable"] = bool(r["unstable"])
    if not args.no_cpu:
        threads = cpu_threads()
        steps = args.cpu_steps or max(2, int(round(12 * 33554432 / (dims[0] * dims[1] * dims[2]))))
        cb = cpu_baseline(r["solver"], steps, threads)
        out["cpu_baseline"] = {"value": round(cb["value"], 2), "unit": "MLUPS", "cores": threads, "kind": "port",
                               "host_cpus": os.cpu_count(),
                               "sample": "%d steps of the same %dx%dx%d sedan domain (the GPU run's flags, wall distances and "
                                         "wheels: Bouzidi links, rolling wheels, belt), FP32 CPU restatement with FP16 store "
                                         "rounding, OpenMP on %d threads, %.1f s" % ((steps,) + dims + (threads, cb["seconds"]))}
        progress("CPU baseline %.1f MLUPS on %d threads" % (cb["value"], threads))
    r["solver"].destroy()
    if not args.no_small and dims != small:
        rs = run_gpu(small, args.steps, args.warmup, args.precision, chunk=args.chunk)
        line = {"mlups": round(rs["mlups"], 1), "ms_per_step": round(rs["wall"] * 1e3 / args.steps, 5),
                "chunk": rs["tune"]["chunk"] if rs["tune"] else args.chunk,
                "kernel_frac": round(rs["achieved_gbs"] / HBM_PEAK_GBS, 4), "avg_launch_ms": round(rs["k_avg_ms"], 5),
                "step_roofline_frac": round(rs["mlups"] * 1e6 * 2 * rs["q"] * rs["sbytes"] / (HBM_PEAK_GBS * 1e9), 4),
                "traffic": traffic_from_profiles("pmc_step_d3q19_%s_%dx%dx%d.txt" % ((args.precision,) + small)),
                "note": "the 318.8 MB ping-pong partly fits the 256 MB memory-side cache: not a roofline figure"}
        if not args.no_physics:
            ph = physics(rs["solver"], rs["body"], rs["info"], 0.08)
            rd = lambda v: round(v, 4) if v == v else None  # noqa: E731
            line["physics"] = {"cd_second_half_5k": rd(ph["cd_half"]), "cl_second_half_5k": rd(ph["cl_half"]),
                               "cd": rd(ph["cd"]), "cd_se": rd(ph["cd_se"]), "cl": rd(ph["cl"]),
                               "converged": ph["converged"], "steps": ph["steps"], "frontal_cells": ph["frontal"],
                               "unstable": ph["unstable"],
                               "window": "cd / cl: the app's readout, last 4 L/U (L = %g cells) from one convective time "
                                         "on, run until the convergence rule holds; *_second_half_5k: steps 2.5k-5k "
                                         "of the configuration's 5k-step run" % ph["L"]}
        out["d3q19_%s_256x128x128" % args.precision] = line
        progress("256x128x128 step and physics")
        plain_ms = rs["k_avg_ms"]
        rs["solver"].destroy()
    elif dims == small:
        plain_ms = r["k_avg_ms"]  # the headline run is the 256x128x128 sedan itself
    if (args.learned_step or not (args.no_learned_step or args.profile_run)) and (dims == small or not args.no_small):
        out["learned_step"] = learned_step_bench(small, max(min(args.steps, 2000) // 4, 50), args.warmup, plain_ms,
                                                 args.precision)
        progress("learned step")
    if not args.no_physics:
        from windtunnel_vis_amd.validation import cylinder_512x128
        from windtunnel_vis_amd import SolverGPU
        c1 = cylinder_512x128(lambda d, p: SolverGPU(d, p))
        rd = lambda w: {"St": round(w["St"], 4) if w["St"] == w["St"] else None, "cd": round(w["cd"], 4),
                        "cycles": w["cycles"], "steps": w["steps"]}
        out["cylinder_512x128"] = {"window_20k": rd(c1["window_20k"]), "developed": rd(c1["developed"]),
                                   "unstable": c1["unstable"], "seconds": round(c1["seconds"], 2)}
        progress("cylinder 512x128")
    if not args.no_large:
        out["shard_128x256x256"] = shard_bench(max(min(args.steps, 2000) // 4, 100), args.warmup, args.precision)
        progress("x-slab shard")
    if not args.no_sweep:
        out["batched_sweep"] = sweep_bench()
        progress("batched sweep")
    if not args.no_closure:
        out["learned_closure"] = closure_bench()
        progress("learned-closure study")
    if not args.no_stress:
        out["precision_stress_1024x512x512"] = precision_stress()
        progress("precision stress")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
