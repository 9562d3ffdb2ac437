// Drives the Node host on a small seeded case and prints populations statistics as JSON, for
// tests/test_gpu_node_host.py to compare with the Python host and the oracle.
import { SolverGPU, Tracers } from '../../windtunnel_vis_amd/node/solver.mjs';

async function main() {
  const dims = { nx: 24, ny: 12, nz: 10 };
  const s = new SolverGPU(dims, { u: 0.06, nu: 0.01, cs: 0.1, ground: 'moving', sides: 'freeslip', spongeNu: 0.04,
    spongeStart: 0.8, spongeIn: 0.15, tauWall: 0.53, emaAlpha: 0.1, precision: process.argv[2] || 'f16' });
  const flags = new Uint32Array(dims.nx * dims.ny * dims.nz);
  for (let z = 3; z < 7; z++) for (let y = 4; y < 8; y++) for (let x = 9; x < 13; x++) flags[x + dims.nx * (y + dims.ny * z)] = 1;
  s.uploadFlags(flags);
  const st = await s.stepAndSample(100);
  const f = await s.readPopulations();
  let sum = 0;
  for (const v of f) sum += v;
  s.derive();
  const e = await s.exportFields(2);
  const esum = (a) => { let t = 0; for (const v of a) t += v; return t; };
  const ex = { dims: e.dims, vel: esum(e.vel), mean: esum(e.mean), vort: esum(e.vort), vort8: Array.from(e.vort.slice(0, 8)) };
  // tracers on the same state
  const pr = s.probe([2.5, 6.0, 5.0, 20.25, 3.5, 7.75, 11.0, 0.2, 9.9]);
  const rake = { a: [1.5, 2.0, 3.0], b: [1.5, 10.0, 3.0], d: [0.0, 0.0, 2.0] };
  const sl = s.streamlines(rake, 6, 12, 0.5, 3, false);
  const tr = new Tracers(s, { count: 64, trail: 4, emitter: 'inlet', maxAge: 50 });
  tr.update(0, { fill: true, advance: false, seed: 1 });
  tr.update(2, { seed: 2 });
  tr.update(2, { seed: 3 });
  const pt = tr.read();
  tr.destroy();
  const tracers = { vel: Array.from(pr.vel), mean: Array.from(pr.mean), sl: Array.from(sl), state: Array.from(pt.state),
    trail: esum(pt.trail), head: pt.head };
  console.log(JSON.stringify({ stats: st, sum, first: Array.from(f.slice(0, 8)), stepCount: s.stepCount, precision: s.precision, ex,
    tracers }));
  s.destroy();
}
main().catch((e) => { console.error(e); process.exit(2); });
