"""Measurement on the host: aerodynamic coefficients, windowed statistics and shedding frequency.

Restates src/analysis/aero.ts:4-52 and src/analysis/strouhal.ts:5-40 (FP64, as the reference's JS).
"""
import math

import numpy as np


def coefficients(fx, fy, fz, U, area, rho=1.0):
    """aero.ts:4-7. In 2D `area` is the frontal height in cells."""
    q = 0.5 * rho * U * U * max(area, 1e-9)
    return dict(cd=fx / q, cl=fy / q, cs=fz / q)


def windowStats(t, v, window):
    """Batch-means mean and standard error over the last `window` time units (aero.ts:33-52)."""
    n = len(t)
    if n < 16:
        return dict(mean=math.nan, se=math.nan, span=0, n=n)
    tEnd = t[n - 1]
    i0 = n - 1
    while i0 > 0 and t[i0 - 1] >= tEnd - window:
        i0 -= 1
    span = tEnd - t[i0]
    B = 8
    sums = [0.0] * B
    cnt = [0] * B
    tot = 0.0
    totN = 0
    for i in range(i0, n):
        b = min(B - 1, int(math.floor(((t[i] - t[i0]) / max(span, 1e-9)) * B)))
        sums[b] += v[i]
        cnt[b] += 1
        tot += v[i]
        totN += 1
    mean = tot / totN
    bm = [s / c for s, c in zip(sums, cnt) if c]
    bm = [x for x in bm if math.isfinite(x)]
    if len(bm) < 4:
        return dict(mean=mean, se=math.nan, span=span, n=totN)
    varB = sum((x - mean) ** 2 for x in bm) / (len(bm) - 1)
    return dict(mean=mean, se=math.sqrt(varB / len(bm)), span=span, n=totN)


def crossingFrequency(t, v):
    """Dominant frequency from upward mean crossings with 15 % hysteresis (strouhal.ts:5-30)."""
    n = len(v)
    if n < 8:
        return dict(f=math.nan, cycles=0, amplitude=math.nan)
    mean = sum(v) / n
    lo, hi = min(v), max(v)
    amp = (hi - lo) / 2
    hyst = amp * 0.15
    ups = []
    armed = v[0] < mean - hyst
    for i in range(1, n):
        if v[i] < mean - hyst:
            armed = True
        if armed and v[i - 1] < mean and v[i] >= mean:
            a = (mean - v[i - 1]) / (v[i] - v[i - 1])
            ups.append(t[i - 1] + a * (t[i] - t[i - 1]))
            armed = False
    if len(ups) < 4:
        return dict(f=math.nan, cycles=max(0, len(ups) - 1), amplitude=amp)
    cycles = len(ups) - 1
    period = (ups[-1] - ups[0]) / cycles
    return dict(f=1 / period, cycles=cycles, amplitude=amp)


def sphereCdRef(re):
    """Schiller-Naumann, strouhal.ts:33-35."""
    return (24 / re) * (1 + 0.15 * re ** 0.687)


def cylinderStRef(re):
    """Williamson fit, strouhal.ts:38-40."""
    return 0.2175 - 5.106 / re


def converged(t, cd, L, U, avg_window=4):
    """App-level convergence rule (app.ts:447-468): window of AVG_WINDOW convective times,
    span >= 3 L/U and SE < max(2 % |mean|, 0.005)."""
    w = windowStats(t, cd, avg_window * L / U)
    ok = (w["span"] >= 3 * L / U) and math.isfinite(w["se"]) and w["se"] < max(0.02 * abs(w["mean"]), 0.005)
    return ok, w


def macroscopic(f, lattice):
    """rho and u from populations of shape (q, n)."""
    f = np.asarray(f, dtype=np.float64)
    rho = f.sum(axis=0)
    u = (lattice.c.T.astype(np.float64) @ f) / rho
    return rho, u


AIR_RHO = 1.225  # kg/m^3 (units.ts:10)


def dragNewtons(cd, speed, areaM2):
    """Real drag force at the chosen speed (aero.ts:10-12)."""
    return 0.5 * AIR_RHO * speed * speed * cd * areaM2


class Ema:
    """Exponential moving average that starts at its first finite sample (aero.ts:15-26)."""

    def __init__(self, k):
        self.k = k
        self.value = math.nan

    def push(self, v):
        if not math.isfinite(v):
            return self.value
        self.value = v if math.isnan(self.value) else self.value + self.k * (v - self.value)
        return self.value

    def reset(self):
        self.value = math.nan


class Series:
    """Fixed-capacity (t, v) series that ignores non-finite values (aero.ts:55-76)."""

    def __init__(self, cap=600):
        self.cap = cap
        self.t = []
        self.v = []

    def push(self, t, v):
        if not math.isfinite(v):
            return
        self.t.append(t)
        self.v.append(v)
        if len(self.t) > self.cap:
            self.t.pop(0)
            self.v.pop(0)

    def clear(self):
        self.t = []
        self.v = []

    def meanSince(self, t0):
        xs = [v for t, v in zip(self.t, self.v) if t >= t0]
        return sum(xs) / len(xs) if xs else math.nan
