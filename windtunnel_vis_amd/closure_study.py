"""Learned-closure study on the truck + trailer (the learned-closure workload).

1. A fine plain-Smagorinsky run (default 512x256x256, FP16) of the truck + trailer SDF (cab
   60x28x24 and trailer 110x30x26 cells, L x H x W on the fine grid, gap 6, ten wheels) at
   u = 0.06, Cs = 0.1, with nu doubled against the coarse run (same Reynolds number); its
   converged C_D is the reference, and snapshots of it are box-filtered 2x onto the coarse grid
   as training records (Dataset.add).
2. The 8-32-32-1 network is trained on the GPU on up to `train_samples` records per epoch.
3. Two coarse runs (default 256x128x128, the same body at half scale): plain Smagorinsky and the
   learned closure. For every run the time to a converged C_D (the app's rule, app.ts:447-468: a
   window of 4 L/U, span >= 3 L/U, SE < max(2 % |mean|, 0.005)) is the wall-clock time from which
   on the rule holds and the window mean stays within 2 % of the fine reference to the end of the
   run; the fine run's own time to its converged C_D is reported beside it.
"""
import math
import time

import numpy as np

from . import closure
from . import geometry as G
from .analysis import coefficients, converged
from .solver import SolverGPU


def truck(dims, scale):
    s = scale
    return G.truck_trailer(dims, cab=(60 * s, 28 * s, 24 * s), trailer=(110 * s, 30 * s, 26 * s), gap=6 * s, corner=4 * s,
                           wheel_r=6 * s, wheel_w=5 * s, ride=3 * s)


def study_params(nu, precision="f16"):
    return dict(u=0.06, nu=nu, cs=0.1, ground="moving", sides="freeslip", spongeNu=0.04, spongeStart=0.84, spongeIn=0.03,
                tauWall=0.53, interp=True, collision="regularized", emaAlpha=0.02, precision=precision)


def make(dims, scale, nu, precision="f16", seed=None):
    """The truck run; with `seed`, the start state carries a seeded velocity perturbation."""
    s = SolverGPU(dims, study_params(nu, precision))
    body = truck(dims, scale)
    s.setWheels(body.wheels)
    info = s.voxelizeSdf(body.prims())
    if seed is not None:
        perturb(s, 0.06, seed)
    return s, body, info


def perturb(s, U, seed, amp=1e-2):
    """Equilibrium start at rho = 1, u = (U, 0, 0) + amp U (uniform noise per component and fluid
    cell, seeded); solid cells at rest. Members of an ensemble differ only in `seed`."""
    rng = np.random.default_rng(seed)
    flags = s.readFlags()
    lat = s.lattice
    n = s.n
    u = np.zeros((3, n), np.float32)
    u[0] = U
    u += (amp * U) * rng.uniform(-1.0, 1.0, (3, n)).astype(np.float32)
    u[:, flags != 0] = 0.0
    c = lat.c.T.astype(np.float32)  # (3, q)
    usq = (u * u).sum(axis=0)
    f = np.empty((lat.q, n), np.float32)
    for i in range(lat.q):
        cu = c[0, i] * u[0] + c[1, i] * u[1] + c[2, i] * u[2]
        f[i] = lat.w[i] * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq)
    s.uploadPopulations(f)


BANDS = (0.02, 0.05, 0.10, 0.15)


def run_case(s, info, L, U, max_steps, chunk=250, target=None, on_chunk=None):
    """Steps `s` in chunks, recording C_D; returns the series and the convergence times (for the
    2 % band and, as a coarser picture, for every band of BANDS)."""
    t, cd, cl = [], [], []
    wall = 0.0
    conv_step = conv_wall = None
    hit_step = hit_wall = None
    band_wall = {b: None for b in BANDS}
    step = 0
    unstable = False
    while step < max_steps:
        t0 = time.perf_counter()
        r = s.stepAndSample(chunk)
        wall += time.perf_counter() - t0
        step += chunk
        if r.unstable:
            unstable = True
            break
        c = coefficients(r.fx, r.fy, r.fz, U, info["frontal"])
        t.append(step)
        cd.append(c["cd"])
        cl.append(c["cl"])
        ok, w = converged(t, cd, L, U)
        if ok and conv_step is None:
            conv_step, conv_wall = step, wall
        # time to the target: the first chunk from which on the rule holds and the window mean
        # stays within 2 % of the target to the end of the run (a transient crossing does not count)
        inside = ok and target is not None and abs(w["mean"] - target) <= 0.02 * abs(target)
        if inside and hit_step is None:
            hit_step, hit_wall = step, wall
        elif not inside:
            hit_step = hit_wall = None
        for b in BANDS:
            if ok and target is not None and abs(w["mean"] - target) <= b * abs(target):
                if band_wall[b] is None:
                    band_wall[b] = wall
            else:
                band_wall[b] = None
        if on_chunk:
            on_chunk(step)
    ok, w = converged(t, cd, L, U)
    return dict(steps=step, wall=wall, cd_mean=w["mean"], cd_se=w["se"], converged=ok, converged_step=conv_step,
                converged_wall=conv_wall, within2_step=hit_step, within2_wall=hit_wall, band_wall=band_wall, unstable=unstable,
                mlups=s.n * step / max(wall, 1e-9) / 1e6)


def study(fine_dims=(512, 256, 256), coarse_dims=(256, 128, 128), nu_coarse=6e-4, fine_steps=30000, coarse_steps=16000,
          snapshot_from=12000, snapshot_every=3000, train_samples=10_000_000, epochs=6, batch=8192, lr=2e-3, seed=20261203,
          precision="f16", smag_floor=1.0, ensemble=3, log=None):
    say = log or (lambda *a: None)
    U = 0.06
    scale_f = fine_dims[0] / 512.0
    coarse, cbody, cinfo = make(coarse_dims, scale_f / 2, nu_coarse, precision)
    fine, fbody, finfo = make(fine_dims, scale_f, 2 * nu_coarse, precision)
    L_f, L_c = fbody.length, cbody.length
    cap = int(coarse.n * (1 + (fine_steps - snapshot_from) // snapshot_every))
    ds = closure.Dataset(cap)
    snaps = []

    def snap(step):
        if step >= snapshot_from and (step - snapshot_from) % snapshot_every == 0:
            snaps.append(ds.add(fine, coarse))

    say("fine run %s" % (fine_dims,))
    rf = run_case(fine, finfo, L_f, U, fine_steps, on_chunk=snap)
    ref = rf["cd_mean"]
    say("fine C_D %.4f (converged %s at step %s), %d records" % (ref, rf["converged"], rf["converged_step"], ds.count))
    fine.destroy()
    net = closure.Network.initial(seed=seed, smag_floor=smag_floor)
    closure.fit_normalization(net, ds)
    t0 = time.perf_counter()
    losses = closure.train(ds, net, epochs=epochs, batch=batch, lr=lr, seed=seed, max_samples=min(train_samples, ds.count))
    t_train = time.perf_counter() - t0
    say("training: %.1f s, loss %s" % (t_train, np.round(losses, 5).tolist()))
    n_records = ds.count
    ds.destroy()
    coarse.reset()
    rp = run_case(coarse, cinfo, L_c, U, coarse_steps, target=ref)
    say("coarse plain C_D %.4f" % rp["cd_mean"])
    coarse.destroy()
    learned, _, linfo = make(coarse_dims, scale_f / 2, nu_coarse, precision)
    closure.attach(learned, net)
    rl = run_case(learned, linfo, L_c, U, coarse_steps, target=ref)
    say("coarse learned C_D %.4f" % rl["cd_mean"])
    learned.destroy()
    # the same two coarse runs from seeded perturbed starts: the spread of the converged C_D over
    # the members is the run-to-run scatter of the measurement itself (a chaotic wake)
    ens = dict(plain=[], learned=[])
    for k in range(ensemble):
        for kind in ("plain", "learned"):
            s, _, info = make(coarse_dims, scale_f / 2, nu_coarse, precision, seed=seed + 1 + k)
            if kind == "learned":
                closure.attach(s, net)
            ens[kind].append(run_case(s, info, L_c, U, coarse_steps, target=ref)["cd_mean"])
            s.destroy()
    say("ensemble C_D plain %s learned %s" % (np.round(ens["plain"], 4).tolist(), np.round(ens["learned"], 4).tolist()))

    def err(r):
        return abs(r["cd_mean"] - ref) / abs(ref) if math.isfinite(r["cd_mean"]) else math.inf

    out = dict(fine_dims=list(fine_dims), coarse_dims=list(coarse_dims), cd_fine=ref, fine=rf, coarse_plain=rp, coarse_learned=rl,
               cd_err_plain=err(rp), cd_err_learned=err(rl), records=n_records, snapshots=snaps, train_seconds=t_train,
               train_loss=losses.tolist(), network=net, ensemble=ens)
    # time to the fine run's C_D (2 %) against the fine run's own time to its converged C_D
    if rl["within2_wall"] is not None and rf["converged_wall"]:
        out["speedup_learned_vs_fine"] = rf["converged_wall"] / rl["within2_wall"]
    if rp["within2_wall"] is not None and rf["converged_wall"]:
        out["speedup_plain_coarse_vs_fine"] = rf["converged_wall"] / rp["within2_wall"]
    return out
