"""Run controller: the physics side of the reference app's run loop (src/app.ts), driving a solver
frame by frame without a browser.

Per frame (app.ts:786-814): the lattice speed ramps toward its target (k = 0.06 per frame); the
viscosity follows a start-up damping target nu (1 + 12 exp(-t / 0.25)) (capped at 0.08) in log
space (0.25 per frame), so the impulsive start's acoustic transient is damped during the first
convective time; the convective clock advances by steps U / L; the EMA rate of the solver's
statistics is 1 - exp(-steps / (3 L / U)), but at least 1 / frames-since-reset. Then the steps run
and the sample is taken. The stability watchdog (app.ts:368-394) reacts to an unstable sample or a
local Mach number above 0.6: it lowers the lattice speed (x0.85 while above 0.55 of nominal, else
raises the viscosity x1.6), always raises the viscosity x1.25, re-maps the flow and restarts it;
it ignores further alarms for 30 frames. Coefficient samples (app.ts:381-420) before the
averaging start (one convective time after a reset, 0.7 after a flow change) are kept out of the
averages; the readout (app.ts:446-466) is the batch-means window of the last 4 convective times.

The reference runs frames asynchronously (stats arrive a frame or two later and are back-dated);
here each frame's sample is taken synchronously, so the back-dating term is zero.
"""
import math

from . import units
from .analysis import Ema, Series, coefficients, dragNewtons, windowStats

AVG_WINDOW = 4
V_MAX_KMH = 300


class RunController:
    def __init__(self, solver, Lcells, frontal, speed_kmh=100.0, lengthM=4.5, nuAir=1.5, reOverride=None, cs=0.1,
                 ground="moving", interp=True, frontalM2=None):
        """solver: a SolverGPU with its body voxelized. Lcells: reference length in cells; frontal:
        voxelized frontal area (cells^2, 3D) or height (cells, 2D). nuAir in 1e-5 m^2/s (the app's
        slider unit). reOverride: simulated Reynolds number instead of the scaled real one."""
        self.solver = solver
        self.Lcells = float(Lcells)
        self.frontal = float(frontal)
        self.speed_kmh = speed_kmh
        self.lengthM = lengthM
        self.nuAir = nuAir
        self.reOverride = reOverride
        self.cs = cs
        self.ground = ground
        self.interp = interp
        self.frontalM2 = frontalM2 if frontalM2 is not None else frontal * (lengthM / Lcells) ** 2
        self.is3D = solver.dims["nz"] > 1
        self.stab = dict(uScale=1.0, nuScale=1.0, events=0)
        self.frames = 0
        self.lastStabFrame = -1000
        self.lastStats = None
        self.lastSampleStep = 0
        self.cdSeries = Series(3000)
        self.raw = dict(cd=Series(20000), cl=Series(20000), cs=Series(20000))
        self.cdAvg, self.clAvg, self.csAvg = Ema(0.05), Ema(0.05), Ema(0.05)
        self.cdInst = math.nan
        self.updateFlow()
        self.resetFlow()

    # -- flow mapping (app.ts:318-339) ----------------------------------------------------------
    def updateFlow(self):
        self.flow = units.mapFlow(self.speed_kmh / 3.6, self.lengthM, self.nuAir * 1e-5, self.Lcells, V_MAX_KMH / 3.6,
                                  0.0006 if self.is3D else 0.0003, self.reOverride, self.stab["uScale"], self.stab["nuScale"])

    def resetFlow(self):
        """app.ts:341-348."""
        self.solver.reset()
        self.clock = 0.0
        self.avgStart = 1.0
        self.cur = dict(U=self.flow["U"], nu=self.flow["nu"])
        self.framesSinceReset = 0
        self.resetCoefficients()

    def resetCoefficients(self):
        self.cdSeries.clear()
        self.cdAvg.reset()
        self.clAvg.reset()
        self.csAvg.reset()
        for r in self.raw.values():
            r.clear()

    def flowChanged(self):
        """After a change of speed, viscosity, Re, floor or LES (app.ts:422-427)."""
        self.updateFlow()
        self.resetCoefficients()
        self.avgStart = max(self.avgStart, self.clock + 0.7)

    # -- one frame ------------------------------------------------------------------------------
    def frame(self, steps):
        self.frames += 1
        self.cur["U"] += (self.flow["U"] - self.cur["U"]) * 0.06
        tConv = self.clock
        nuTarget = min(self.flow["nu"] * (1 + 12 * math.exp(-tConv / 0.25)), 0.08)
        self.cur["nu"] = math.exp(math.log(self.cur["nu"]) + (math.log(nuTarget) - math.log(self.cur["nu"])) * 0.25)
        self.clock += steps * self.cur["U"] / self.Lcells
        avgTime = 3 * self.Lcells / max(self.cur["U"], 1e-4)
        ema = 1 - math.exp(-max(steps, 1) / avgTime)
        self.framesSinceReset += 1
        ema = max(ema, 1 / self.framesSinceReset)
        self.solver.setParams(dict(u=self.cur["U"], nu=self.cur["nu"], cs=self.cs, ground=self.ground,
                                   emaAlpha=ema if steps > 0 else 0.0, interp=self.interp))
        st = self.solver.stepAndSample(steps)
        self.onStats(st)
        return st

    # -- stability watchdog (app.ts:368-377) ----------------------------------------------------
    def handleInstability(self, reason):
        self.lastStabFrame = self.frames
        self.stab["events"] += 1
        if self.stab["uScale"] > 0.55:
            self.stab["uScale"] *= 0.85
        else:
            self.stab["nuScale"] *= 1.6
        self.stab["nuScale"] *= 1.25
        self.lastReason = reason
        self.updateFlow()
        self.resetFlow()

    # -- samples (app.ts:381-420) ---------------------------------------------------------------
    def onStats(self, st):
        if st.step == 0:
            return
        if st.step > self.solver.stepCount:
            return
        self.lastStats = st
        mach = st.maxU * math.sqrt(3)
        if (st.unstable or mach > 0.6) and self.frames - self.lastStabFrame < 30:
            return
        if st.unstable:
            self.handleInstability("Ma = %.2f" % mach if math.isfinite(st.fx) else "NaN")
            return
        if mach > 0.6:
            self.handleInstability("local Mach %.2f > 0.6" % mach)
            return
        U = self.cur["U"]
        c = coefficients(st.fx, st.fy, st.fz, U, max(self.frontal, 1), st.rhoRef)
        tConv = self.clock - max(0, self.solver.stepCount - st.step) * U / self.Lcells
        dt = st.step - self.lastSampleStep if self.lastSampleStep else 0
        self.lastSampleStep = st.step
        k = 1 - math.exp(-(dt * U) / (1.5 * self.Lcells)) if dt > 0 else 0.05
        self.cdAvg.k = self.clAvg.k = self.csAvg.k = min(max(k, 0.002), 1)
        self.cdInst = c["cd"]
        if tConv < self.avgStart:
            self.cdSeries.push(tConv, c["cd"])
            return
        self.cdAvg.push(c["cd"])
        self.clAvg.push(c["cl"])
        self.csAvg.push(c["cs"])
        self.raw["cd"].push(tConv, c["cd"])
        self.raw["cl"].push(tConv, c["cl"])
        self.raw["cs"].push(tConv, c["cs"])
        self.cdSeries.push(tConv, c["cd"])

    @property
    def settleLeft(self):
        return max(0.0, self.avgStart - self.clock)

    def readout(self):
        """app.ts:446-466 (the host-side figures; frame rate and MLUPS are the caller's)."""
        w = lambda s: windowStats(s.t, s.v, AVG_WINDOW)
        sd, sl, ss = w(self.raw["cd"]), w(self.raw["cl"]), w(self.raw["cs"])
        pick = lambda st, ema: st["mean"] if math.isfinite(st["mean"]) else ema.value
        cd = pick(sd, self.cdAvg)
        speed = self.speed_kmh / 3.6
        d = self.solver.dims
        blockage = self.frontal / (d["ny"] * d["nz"] if self.is3D else d["ny"])
        conv = sd["span"] >= AVG_WINDOW * 0.75 and math.isfinite(sd["se"]) and sd["se"] < max(0.02 * abs(sd["mean"]), 0.005)
        return dict(cd=cd, cl=pick(sl, self.clAvg), cs=pick(ss, self.csAvg), cdInst=self.cdInst, cdSE=sd["se"],
                    avgSpan=sd["span"], converged=conv, cdCorrected=cd * (1 - blockage) ** 2, frontalM2=self.frontalM2,
                    dragN=dragNewtons(cd, speed, self.frontalM2), flow=dict(self.flow), convTime=self.clock,
                    maxMach=self.lastStats.maxU * math.sqrt(3) if self.lastStats else 0.0, blockage=blockage)
