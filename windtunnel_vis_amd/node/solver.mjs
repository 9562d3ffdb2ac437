// Node host of the MI355X solver: the reference's SolverGPU surface (src/solver/SolverGPU.ts) and
// its LbmSolver / SolverFactory contract (src/backend/types.ts:7-18), over the N-API addon that
// binds libwt.so. Types: solver.d.ts.
import { createRequire } from 'module';
import { dirname, join } from 'path';
import { fileURLToPath } from 'url';

const require = createRequire(import.meta.url);
const here = dirname(fileURLToPath(import.meta.url));
const native = require(join(here, '..', '_lib', 'wt_addon.node'));

const GROUND = { freeslip: 0, noslip: 1, moving: 2 };
const SIDES = { freeslip: 0, periodic: 1 };
const PRECISION = { f32: 0, f16: 1, bf16: 2 };
const PRECISION_NAME = ['f32', 'f16', 'bf16'];
const COLLISION = { regularized: 0, bgk: 1 };
const CLOSURE = { smagorinsky: 0, mlp: 1 };

/** D2Q9 / D3Q19 velocity sets (src/solver/lattice.ts:39-60). */
export const D2Q9 = {
  name: 'D2Q9', dim: 2, q: 9,
  c: [[0, 0, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0], [1, 1, 0], [-1, 1, 0], [-1, -1, 0], [1, -1, 0]],
  w: [4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36],
};
export const D3Q19 = {
  name: 'D3Q19', dim: 3, q: 19,
  c: [[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0], [-1, -1, 0], [1, 0, 1],
    [-1, 0, -1], [0, 1, 1], [0, -1, -1], [1, -1, 0], [-1, 1, 0], [1, 0, -1], [-1, 0, 1], [0, 1, -1], [0, -1, 1]],
  w: [1 / 3, ...Array(6).fill(1 / 18), ...Array(12).fill(1 / 36)],
};

/** SolverParams (src/solver/types.ts:3-29) -> the addon's numeric form. Only present keys are sent. */
function encodeParams(p) {
  const o = {};
  const num = ['u', 'nu', 'cs', 'spongeNu', 'spongeStart', 'spongeIn', 'tauWall', 'emaAlpha'];
  for (const k of num) if (p[k] !== undefined && p[k] !== null) o[k] = p[k];
  if (p.ground !== undefined) o.ground = GROUND[p.ground];
  if (p.sides !== undefined) o.sides = SIDES[p.sides];
  if (p.interp !== undefined) o.interp = p.interp === false ? 0 : 1;
  if (p.collision !== undefined) o.collision = COLLISION[p.collision];
  if (p.closure !== undefined) o.closure = CLOSURE[p.closure];
  if (p.precision !== undefined) o.precision = PRECISION[p.precision];
  for (const [k, v] of Object.entries(o)) if (v === undefined) throw new TypeError(`bad value for params.${k}`);
  return o;
}

function wheelArray(list) {
  const n = Math.min(list.length, 16);
  const a = new Float32Array(8 * n);
  for (let k = 0; k < n; k++) {
    const w = list[k];
    a.set([w.c[0], w.c[1], w.c[2], w.axis[0], w.axis[1], w.axis[2], w.r, w.hw], 8 * k);
  }
  return a;
}

/**
 * GPU lattice Boltzmann solver (D2Q9 when nz == 1, D3Q19 otherwise): SolverGPU.ts:12-444.
 * The constructor takes (dims, params) — the device is the process's HIP device.
 */
export class SolverGPU {
  constructor(dims, params) {
    this.dims = { nx: dims.nx, ny: dims.ny, nz: dims.nz };
    this.is3D = dims.nz > 1;
    this.lattice = this.is3D ? D3Q19 : D2Q9;
    this.n = dims.nx * dims.ny * dims.nz;
    this.p = { tauWall: 0.5, spongeIn: 0, interp: true, collision: 'regularized', precision: 'f32', ...params };
    this.h = native.create(this.dims, encodeParams(this.p));
    const info = native.info(this.h);
    this.precision = PRECISION_NAME[info.precision];
    /** populations stored as packed half-width deviation pairs */
    this.packed = this.precision !== 'f32';
    this.latest = null;
    this.pendingReads = [];
    this.pollTimer = null;
  }

  get stepCount() { return native.info(this.h).step; }
  get params_() { return this.p; }
  /** Current tau0 (molecular relaxation time), SolverGPU.ts:151. */
  get tau0() { return 3 * this.p.nu + 0.5; }

  setParams(p) {
    Object.assign(this.p, p);
    native.setParams(this.h, encodeParams(p));
  }

  reset() { native.reset(this.h); }
  clearStats() { native.clearStats(this.h); }

  /** Record `steps` steps (asynchronous on the GPU), SolverGPU.ts:204-222. */
  encodeSteps(steps) { native.step(this.h, steps); }

  /** Run `steps` steps, then read the force / statistics result (SolverGPU.ts:292-309). */
  stepAndSample(steps) {
    // the steps, reduction and read are enqueued now (call order is the GPU order, as the
    // reference's submit); only the wait for the read runs off the JS thread
    return native.stepAndSample(this.h, steps).then((s) => {
      if (!this.latest || s.step >= this.latest.step) this.latest = s;
      return s;
    });
  }

  /**
   * Macro fields, statistics and force reduction (SolverGPU.ts:225-261): with `derived` the derived
   * pass runs too and a read of the statistics is queued (up to four in flight; a post that finds
   * them all busy is not read back, as in the reference). Returns at once.
   */
  encodePost(derived = true) {
    if (native.encodePost(this.h, !!derived)) this.pendingReads.push({ cb: null, stats: null });
  }

  /**
   * Call after encodePost (SolverGPU.ts:264-287): the reads queued since the last flush deliver
   * their statistics to `onStats` as they complete, in any later turn of the event loop, and
   * `latest` keeps the newest by step. Never blocks: completed reads are polled. A read that
   * completes before its flush (posted after the previous one) is held until that flush, as the
   * reference maps a read only when it is flushed.
   */
  flushReads(onStats) {
    for (const r of this.pendingReads) if (!r.cb) r.cb = onStats || (() => {});
    this._pollReads();
  }

  _pollReads() {
    if (this.pollTimer || !this.h || this.pendingReads.length === 0) return;
    const poll = () => {
      this.pollTimer = null;
      if (!this.h) return;
      // reads complete in posting order: hand the completed ones to the oldest open entries
      const { stats } = native.flushReads(this.h);
      let k = 0;
      for (const r of this.pendingReads) {
        if (k >= stats.length) break;
        if (!r.stats) r.stats = stats[k++];
      }
      // deliver the completed, flushed prefix (flushed entries precede the ones posted after)
      while (this.pendingReads.length && this.pendingReads[0].stats && this.pendingReads[0].cb) {
        const r = this.pendingReads.shift();
        if (!this.latest || r.stats.step >= this.latest.step) this.latest = r.stats;
        r.cb(r.stats);
      }
      if (this.pendingReads.some((r) => r.cb)) this.pollTimer = setTimeout(poll, 1);
    };
    this.pollTimer = setTimeout(poll, 0);
  }

  /**
   * device.queue.onSubmittedWorkDone() for this solver's queue (the handle's stream): a Promise that
   * resolves once the GPU has finished the work queued on the solver so far. The reference's frame
   * loop counts frames in flight with it (backend/webgpu.ts:92-94, :120). Polled, never blocks; it
   * also resolves when the solver is destroyed.
   */
  onSubmittedWorkDone() {
    const t = native.fence(this.h);
    return new Promise((resolve) => {
      const poll = () => {
        if (!this.h || native.fenceDone(this.h, t)) resolve();
        else setTimeout(poll, 1);
      };
      poll();
    });
  }

  uploadFlags(flags) { native.uploadFlags(this.h, flags); }
  async readFlags() { return native.readFlags(this.h); }
  /** Distances to the wall (cells, < 0 = unknown) for interpolated bounce-back, SolverGPU.ts:312. */
  uploadSdf(d) { native.uploadSdf(this.h, d); }
  /** Rotating wheels in grid space (at most 16), SolverGPU.ts:325-335. */
  setWheels(list) { native.setWheels(this.h, wheelArray(list)); }
  /** Current populations, unpacked to full f_i (q * n floats). */
  async readPopulations() { return native.readPopulations(this.h); }
  uploadPopulations(f) { native.uploadPopulations(this.h, f); }

  /** Analytic-SDF body (list of primitives, see geometry.mjs). */
  voxelizeSdf(prims) { return native.voxelizeSdf(this.h, prims); }
  /** VoxelizerGPU.voxelize(solver, mesh, matrix, wheels), VoxelizerGPU.ts:48-151. */
  voxelizeMesh(mesh, matrix, wheels = []) {
    return native.voxelizeMesh(this.h, Float32Array.from(mesh.positions), Uint32Array.from(mesh.indices),
      Float32Array.from(matrix), wheelArray(wheels));
  }

  /** Derived pass of encodePost(derived = true) (SolverGPU.ts:242-246) over the last sample's fields. */
  derive() { native.derive(this.h); }
  /** SolverGPU.exportFields (SolverGPU.ts:374-425): fields sub-sampled by `stride`, FieldExport. */
  async exportFields(stride) {
    const r = native.exportFields(this.h, stride);
    return { dims: r.dims, stride, vel: r.vel, mean: r.mean, vort: r.vort };
  }

  /**
   * PROBES_WGSL (render/tracers.ts:109-125): the velocity record and the mean record at points in
   * cell units (a flat array of x, y, z triples); returns { vel, mean }, 4 floats per point
   * (ux, uy, uz, rho) sampled linearly with clamp to edge.
   */
  probe(points) {
    const k = points.length / 3;
    const pos = new Float32Array(4 * k);
    for (let i = 0; i < k; i++) for (let a = 0; a < 3; a++) pos[4 * i + a] = points[3 * i + a];
    const out = native.probe(this.h, pos);
    const vel = new Float32Array(4 * k), mean = new Float32Array(4 * k);
    for (let i = 0; i < k; i++) for (let a = 0; a < 4; a++) {
      vel[4 * i + a] = out[8 * i + a];
      mean[4 * i + a] = out[8 * i + 4 + a];
    }
    return { vel, mean };
  }

  /**
   * STREAMLINES_WGSL (render/tracers.ts:152-200): `seeds` lines seeded on the rake
   * ({a, b, d}: 3-vectors; perLine seeds per rake line), `points` points each, step h along the
   * normalised velocity (midpoint rule); Float32Array of seeds x points x (x, y, z, speed or -1).
   */
  streamlines(rake, seeds, points, h = 0.5, perLine = seeds, useMean = false) {
    return native.streamlines(this.h, rakeArray(rake), seeds, perLine, points, h, !!useMean);
  }

  /**
   * destroy(), SolverGPU.ts:432-437. Safe while a stepAndSample is in flight: the addon frees the
   * solver when that sample completes (its promise still settles); pending reads are dropped.
   */
  destroy() {
    if (this.pollTimer) clearTimeout(this.pollTimer);
    this.pollTimer = null;
    this.pendingReads = [];
    if (this.h) native.destroy(this.h);
    this.h = null;
  }

  static bytesFor(dims) { return native.bytesFor(dims); }
}

/** SolverFactory (src/backend/types.ts:18). */
export const solverFactory = (dims, p) => new SolverGPU(dims, p);

/** A rake {a, b, d} (3-vectors, d = offset between parallel lines) as the addon's Float32Array(12). */
function rakeArray(r = {}) {
  const out = new Float32Array(12);
  const put = (v, o) => { if (v) for (let k = 0; k < 3; k++) out[o + k] = v[k] || 0; };
  put(r.a, 0);
  put(r.b, 4);
  put(r.d, 8);
  return out;
}

/**
 * PARTICLES_WGSL (render/tracers.ts:3-106): particle tracers advected on a solver's velocity field
 * with a trail ring buffer. emitter 'inlet' spawns on the inlet plane (x in [1, 3)), 'rake' from
 * `nozzles` nozzles on each of `lines` rake lines. Destroy it before its solver.
 */
export class Tracers {
  constructor(solver, { count, trail = 16, emitter = 'inlet', nozzles = 16, lines = 1, maxAge = 400, rake } = {}) {
    this.solver = solver;
    this.count = count;
    this.trail = trail;
    this.frameSeed = 0;
    this.h = native.particlesCreate(solver.h, {
      count, trail, emitter: emitter === 'inlet' ? 0 : 1, nozzles, lines, maxAge, rake: rakeArray(rake),
    });
  }

  /** One frame (RendererGPU.ts:505-527): `steps` lattice steps of advection; fill respawns all. */
  update(steps, { advance = true, fill = false, seed } = {}) {
    if (seed === undefined) seed = this.frameSeed++;
    native.particlesUpdate(this.h, steps, seed >>> 0, !!fill, !!advance);
  }

  /** { state: count x (x, y, z, age), trail: trail x count x 4 (slot-major), head }. */
  read() { return native.particlesRead(this.h); }

  destroy() {
    if (this.h) native.particlesDestroy(this.h);
    this.h = null;
  }
}
