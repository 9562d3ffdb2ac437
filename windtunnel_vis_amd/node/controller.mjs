// Run controller of the Node host: the physics side of the reference app's run loop (src/app.ts)
// and the lattice <-> SI mapping (src/analysis/units.ts), driving a SolverGPU frame by frame
// without a browser. Same rules as windtunnel_vis_amd/controller.py: speed ramp 6 % per frame,
// start-up viscosity damping, convective clock, EMA-rate schedule, stability watchdog with a
// 30-frame hold-off, averages from one convective time after a reset, batch-means readout.
import { coefficients, windowStats } from './analysis.mjs';

export const AIR_RHO = 1.225;
export const AIR_NU = 1.5e-5;
export const CS = 1 / Math.sqrt(3);
const AVG_WINDOW = 4;
const V_MAX_KMH = 300;

/** Inlet speed in cells/step, 0.03 at rest up to 0.08 at vMax (units.ts:50-53). */
export function latticeSpeed(speed, vMax) {
  const s = Math.min(Math.max(speed / vMax, 0), 1);
  return 0.03 + (0.08 - 0.03) * s;
}

/** units.ts:62-74. */
export function mapFlow(s) {
  const reReal = (s.speed * s.lengthM) / s.nuAir;
  const U = latticeSpeed(s.speed, s.vMax) * s.uScale;
  const reMaxGrid = (0.08 * s.Lcells) / s.nuMin;
  const reRealMax = (s.vMax * s.lengthM) / AIR_NU;
  let reSim = s.reOverride && s.reOverride > 0 ? s.reOverride : reReal * (reMaxGrid / reRealMax);
  let nu = (U * s.Lcells) / reSim;
  nu = Math.max(nu * s.nuScale, s.nuMin * 0.5);
  reSim = (U * s.Lcells) / nu;
  const dx = s.lengthM / s.Lcells;
  const dt = s.speed > 0 ? (U / s.speed) * dx : 0;
  return { reReal, reSim, U, nu, tau: 3 * nu + 0.5, mach: U / CS, dx, dt };
}

export function dragNewtons(cd, speed, areaM2) { return 0.5 * AIR_RHO * speed * speed * cd * areaM2; }

export class Ema {
  constructor(k) { this.k = k; this.value = NaN; }
  push(v) {
    if (!isFinite(v)) return this.value;
    this.value = isNaN(this.value) ? v : this.value + this.k * (v - this.value);
    return this.value;
  }
  reset() { this.value = NaN; }
}

export class Series {
  constructor(cap = 600) { this.cap = cap; this.t = []; this.v = []; }
  push(t, v) {
    if (!isFinite(v)) return;
    this.t.push(t); this.v.push(v);
    if (this.t.length > this.cap) { this.t.shift(); this.v.shift(); }
  }
  clear() { this.t = []; this.v = []; }
}

export class RunController {
  /**
   * solver: SolverGPU (or any object with dims, stepCount, reset, setParams, stepAndSample).
   * o: {Lcells, frontal, speedKmh = 100, lengthM = 4.5, nuAir = 1.5 (1e-5 m^2/s), reOverride, cs = 0.1,
   *     ground = 'moving', interp = true, frontalM2}
   */
  constructor(solver, o) {
    this.solver = solver;
    this.o = { speedKmh: 100, lengthM: 4.5, nuAir: 1.5, reOverride: null, cs: 0.1, ground: 'moving', interp: true, ...o };
    this.frontalM2 = this.o.frontalM2 != null ? this.o.frontalM2 : this.o.frontal * (this.o.lengthM / this.o.Lcells) ** 2;
    this.is3D = solver.dims.nz > 1;
    this.stab = { uScale: 1, nuScale: 1, events: 0 };
    this.frames = 0;
    this.lastStabFrame = -1000;
    this.lastStats = null;
    this.lastSampleStep = 0;
    this.cdSeries = new Series(3000);
    this.raw = { cd: new Series(20000), cl: new Series(20000), cs: new Series(20000) };
    this.cdAvg = new Ema(0.05); this.clAvg = new Ema(0.05); this.csAvg = new Ema(0.05);
    this.cdInst = NaN;
    this.updateFlow();
    this.resetFlow();
  }

  updateFlow() {
    const o = this.o;
    this.flow = mapFlow({ speed: o.speedKmh / 3.6, lengthM: o.lengthM, nuAir: o.nuAir * 1e-5, Lcells: o.Lcells,
      vMax: V_MAX_KMH / 3.6, nuMin: this.is3D ? 0.0006 : 0.0003, reOverride: o.reOverride,
      uScale: this.stab.uScale, nuScale: this.stab.nuScale });
  }

  resetFlow() {
    this.solver.reset();
    this.clock = 0;
    this.avgStart = 1;
    this.cur = { U: this.flow.U, nu: this.flow.nu };
    this.framesSinceReset = 0;
    this.resetCoefficients();
  }

  resetCoefficients() {
    this.cdSeries.clear();
    this.cdAvg.reset(); this.clAvg.reset(); this.csAvg.reset();
    for (const r of Object.values(this.raw)) r.clear();
  }

  flowChanged() {
    this.updateFlow();
    this.resetCoefficients();
    this.avgStart = Math.max(this.avgStart, this.clock + 0.7);
  }

  async frame(steps) {
    this.frames++;
    const L = this.o.Lcells;
    this.cur.U += (this.flow.U - this.cur.U) * 0.06;
    const nuTarget = Math.min(this.flow.nu * (1 + 12 * Math.exp(-this.clock / 0.25)), 0.08);
    this.cur.nu = Math.exp(Math.log(this.cur.nu) + (Math.log(nuTarget) - Math.log(this.cur.nu)) * 0.25);
    this.clock += (steps * this.cur.U) / L;
    const avgTime = (3 * L) / Math.max(this.cur.U, 1e-4);
    let ema = 1 - Math.exp(-Math.max(steps, 1) / avgTime);
    this.framesSinceReset++;
    ema = Math.max(ema, 1 / this.framesSinceReset);
    this.solver.setParams({ u: this.cur.U, nu: this.cur.nu, cs: this.o.cs, ground: this.o.ground,
      emaAlpha: steps > 0 ? ema : 0, interp: this.o.interp });
    const st = await this.solver.stepAndSample(steps);
    this.onStats(st);
    return st;
  }

  handleInstability(reason) {
    this.lastStabFrame = this.frames;
    this.stab.events++;
    if (this.stab.uScale > 0.55) this.stab.uScale *= 0.85;
    else this.stab.nuScale *= 1.6;
    this.stab.nuScale *= 1.25;
    this.lastReason = reason;
    this.updateFlow();
    this.resetFlow();
  }

  onStats(st) {
    if (st.step === 0) return;
    if (st.step > this.solver.stepCount) return;
    this.lastStats = st;
    const mach = st.maxU * Math.sqrt(3);
    if ((st.unstable || mach > 0.6) && this.frames - this.lastStabFrame < 30) return;
    if (st.unstable) { this.handleInstability(isFinite(st.fx) ? `Ma = ${mach.toFixed(2)}` : 'NaN'); return; }
    if (mach > 0.6) { this.handleInstability(`local Mach ${mach.toFixed(2)} > 0.6`); return; }
    const U = this.cur.U, L = this.o.Lcells;
    const c = coefficients(st.fx, st.fy, st.fz, U, Math.max(this.o.frontal, 1), st.rhoRef);
    const tConv = this.clock - (Math.max(0, this.solver.stepCount - st.step) * U) / L;
    const dt = this.lastSampleStep ? st.step - this.lastSampleStep : 0;
    this.lastSampleStep = st.step;
    const k = dt > 0 ? 1 - Math.exp(-(dt * U) / (1.5 * L)) : 0.05;
    this.cdAvg.k = this.clAvg.k = this.csAvg.k = Math.min(Math.max(k, 0.002), 1);
    this.cdInst = c.cd;
    if (tConv < this.avgStart) { this.cdSeries.push(tConv, c.cd); return; }
    this.cdAvg.push(c.cd); this.clAvg.push(c.cl); this.csAvg.push(c.cs);
    this.raw.cd.push(tConv, c.cd); this.raw.cl.push(tConv, c.cl); this.raw.cs.push(tConv, c.cs);
    this.cdSeries.push(tConv, c.cd);
  }

  get settleLeft() { return Math.max(0, this.avgStart - this.clock); }

  readout() {
    const w = (s) => windowStats(s.t, s.v, AVG_WINDOW);
    const sd = w(this.raw.cd), sl = w(this.raw.cl), ss = w(this.raw.cs);
    const pick = (st, ema) => (isFinite(st.mean) ? st.mean : ema.value);
    const cd = pick(sd, this.cdAvg);
    const d = this.solver.dims;
    const blockage = this.o.frontal / (this.is3D ? d.ny * d.nz : d.ny);
    const converged = sd.span >= AVG_WINDOW * 0.75 && isFinite(sd.se) && sd.se < Math.max(0.02 * Math.abs(sd.mean), 0.005);
    return { cd, cl: pick(sl, this.clAvg), cs: pick(ss, this.csAvg), cdInst: this.cdInst, cdSE: sd.se, avgSpan: sd.span, converged,
      cdCorrected: cd * (1 - blockage) ** 2, frontalM2: this.frontalM2, dragN: dragNewtons(cd, this.o.speedKmh / 3.6, this.frontalM2),
      flow: { ...this.flow }, convTime: this.clock, maxMach: this.lastStats ? this.lastStats.maxU * Math.sqrt(3) : 0, blockage };
  }
}
