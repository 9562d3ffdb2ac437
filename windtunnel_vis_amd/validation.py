"""Validation cases with the reference's setups and pass bands (src/analysis/validation.ts:60-236),
run on a SolverFactory (the MI355X solver or the CPU oracle)."""
import math
import time

import numpy as np

from .analysis import coefficients, crossingFrequency, cylinderStRef, sphereCdRef
from .geometry import round_flags


def freestream(make):
    """validation.ts:61-88: empty D2Q9 256x96 tunnel, U = 0.08; |maxU/U - 1| < 1 %, |rho_ref - 1| < 1 %."""
    t0 = time.time()
    d = (256, 96, 1)
    U = 0.08
    s = make(d, dict(u=U, nu=0.01, cs=0.1, ground="freeslip", sides="freeslip", spongeNu=0.05, spongeStart=0.85, emaAlpha=0.02))
    s.uploadFlags(np.zeros(d[0] * d[1], np.uint32))
    st = s.stepAndSample(1)
    for _ in range(20):
        st = s.stepAndSample(50)
    s.destroy()
    err = abs(st.maxU - U) / U
    drho = abs(st.rhoRef - 1)
    return dict(id="freestream", passed=err < 0.01 and drho < 0.01 and not st.unstable, seconds=time.time() - t0,
                metrics={"maxU/U": st.maxU / U, "rhoRef": st.rhoRef})


def cylinder(make, scale=1.0):
    """validation.ts:89-136: D2Q9, D = 24 k, 25D x 10D, U = 0.06, Re = 100, Bouzidi; St in [0.16, 0.20], >= 4 cycles."""
    t0 = time.time()
    D = round(24 * scale)
    d = (round(25 * D), 10 * D, 1)
    U, Re = 0.06, 100
    nu = U * D / Re
    s = make(d, dict(u=U, nu=nu, cs=0, ground="freeslip", sides="freeslip", spongeNu=0.03, spongeStart=0.85, emaAlpha=0.0))
    flags, frontal, sdf = round_flags(d, (6 * D, d[1] / 2 + 0.37, 0), D / 2, True)
    s.uploadFlags(flags)
    s.uploadSdf(sdf)
    period = D / (0.17 * U)
    total = round(period * 16)
    measure_from = round(period * 6)
    chunk = max(5, round(period / 40))
    tt, cl, cd = [], [], []
    for step in range(0, total, chunk):
        st = s.stepAndSample(chunk)
        if st.unstable:
            break
        c = coefficients(st.fx, st.fy, 0, U, D)
        if step >= measure_from:
            tt.append(st.step)
            cl.append(c["cl"])
            cd.append(c["cd"])
    s.destroy()
    cf = crossingFrequency(tt, cl)
    St = cf["f"] * D / U
    cd_mean = sum(cd) / max(len(cd), 1)
    return dict(id="cylinder", passed=bool(0.16 <= St <= 0.2 and cf["cycles"] >= 4), seconds=time.time() - t0,
                metrics={"St": St, "St_ref": cylinderStRef(Re), "cd": cd_mean, "cl_amp": cf["amplitude"], "cycles": cf["cycles"],
                         "grid": d})


def sphere(make, scale=1.0):
    """validation.ts:137-184: D3Q19, D = 20 k, 9.6D x 4.8D x 4.8D, U = 0.05, Re = 100; |C_D / SN - 1| < 20 %."""
    t0 = time.time()
    D = round(20 * scale)
    d = (round(9.6 * D), round(4.8 * D), round(4.8 * D))
    U, Re = 0.05, 100
    nu = U * D / Re
    s = make(d, dict(u=U, nu=nu, cs=0, ground="freeslip", sides="freeslip", spongeNu=0.03, spongeStart=0.85, emaAlpha=0))
    flags, frontal, sdf = round_flags(d, (2.8 * D, d[1] / 2 + 0.37, d[2] / 2 + 0.21), D / 2, False)
    s.uploadFlags(flags)
    s.uploadSdf(sdf)
    flow_through = d[0] / U
    total = round(flow_through * 2.2)
    measure_from = round(flow_through * 1.4)
    tt, cds = [], []
    for step in range(0, total, 50):
        st = s.stepAndSample(50)
        if st.unstable:
            break
        tt.append(st.step)
        cds.append(coefficients(st.fx, st.fy, st.fz, U, frontal)["cd"])
    s.destroy()
    sel = [c for t, c in zip(tt, cds) if t >= measure_from]
    cd = sum(sel) / max(len(sel), 1)
    ref = sphereCdRef(Re)
    err = (cd - ref) / ref
    return dict(id="sphere", passed=abs(err) < 0.2, seconds=time.time() - t0,
                metrics={"cd": cd, "cd_ref": ref, "err": err, "grid": d, "blockage": frontal / (d[1] * d[2])})


CASES = {"freestream": freestream, "cylinder": cylinder, "sphere": sphere}
