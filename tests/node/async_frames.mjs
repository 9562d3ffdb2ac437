// The reference's interactive frame pattern through the Node host (backend/webgpu.ts:96-113):
// encodeSteps + encodePost + flushReads every frame, with reads still pending when the next frame is
// encoded; then calls made while a stepAndSample is in flight: encodeSteps right after an
// un-awaited stepAndSample, and destroy() with the sample still pending. Prints one JSON line.
import { SolverGPU } from '../../windtunnel_vis_amd/node/solver.mjs';

const dims = { nx: 24, ny: 12, nz: 10 };
const params = { u: 0.06, nu: 0.01, cs: 0.1, ground: 'moving', spongeNu: 0.04, spongeStart: 0.8, spongeIn: 0.15,
  tauWall: 0.53, emaAlpha: 0.1, precision: 'f16' };
const flags = new Uint32Array(dims.nx * dims.ny * dims.nz);
for (let z = 3; z < 7; z++) for (let y = 4; y < 8; y++) for (let x = 9; x < 13; x++) flags[x + dims.nx * (y + dims.ny * z)] = 1;

// resolves once every queued read of `s` has been delivered
function drain(s) {
  return new Promise((res, rej) => {
    const t0 = Date.now();
    const t = setInterval(() => {
      if (s.pendingReads.length === 0) { clearInterval(t); res(); }
      if (Date.now() - t0 > 10000) { clearInterval(t); rej(new Error('reads did not complete')); }
    }, 2);
  });
}

async function main() {
  const out = {};
  // 1. frames: 10 steps, post, flush; the callbacks arrive later, in order
  {
    const s = new SolverGPU(dims, params);
    s.uploadFlags(flags);
    const got = [];
    // six frames back to back: the first four posts find a free slot; later ones may find all four
    // still waiting for the poll and are then not read back (as in the reference)
    for (let f = 0; f < 6; f++) {
      s.encodeSteps(10);
      s.encodePost(true);
      s.flushReads((st) => got.push(st));
    }
    await drain(s);
    out.burst = got.map((st) => st.step);
    // one more frame once the slots are free: its read is always delivered
    s.encodePost(true);
    s.flushReads((st) => got.push(st));
    await drain(s);
    out.frames = got.map((st) => ({ step: st.step, fx: st.fx, maxU: st.maxU }));
    out.latest = s.latest.step;
    // the same state sampled synchronously: the last delivered read equals it
    const sync = await s.stepAndSample(0);
    out.syncStep = sync.step;
    out.syncFx = sync.fx;
    s.destroy();
  }
  // 2. a read posted in one task and flushed in a later one, after the earlier flush's polls have
  // run (and found it complete): it goes to the later flush's callback
  {
    const s = new SolverGPU(dims, params);
    s.uploadFlags(flags);
    const got = [];
    s.encodeSteps(10);
    s.encodePost(true);
    s.flushReads((st) => got.push(['first', st.step]));
    s.encodeSteps(10);
    s.encodePost(true);
    await new Promise((res) => setTimeout(res, 300));
    s.flushReads((st) => got.push(['second', st.step]));
    await drain(s);
    out.late = got;
    s.destroy();
  }
  // 3. the reference's throttled frame loop (backend/webgpu.ts:92-131): a frame is submitted only
  // while fewer than two are in flight, counted with onSubmittedWorkDone, and setParams ramps the
  // speed every frame (app.ts:792-810); setParams does not wait for the queue
  {
    const s = new SolverGPU(dims, params);
    s.uploadFlags(flags);
    let inFlight = 0, maxInFlight = 0, frames = 0;
    const got = [];
    await new Promise((resolve, reject) => {
      const t0 = Date.now();
      const tick = () => {
        if (frames >= 40) {
          if (inFlight === 0) { resolve(); return; }
        } else if (inFlight < 2) {
          s.setParams({ u: 0.06 * (0.5 + 0.5 * frames / 40) });
          s.encodeSteps(5);
          s.encodePost(true);
          inFlight++;
          frames++;
          maxInFlight = Math.max(maxInFlight, inFlight);
          s.flushReads((st) => got.push(st.step));
          s.onSubmittedWorkDone().then(() => { inFlight--; });
        }
        if (Date.now() - t0 > 20000) { reject(new Error('frames did not complete')); return; }
        setTimeout(tick, 0);
      };
      tick();
    });
    await drain(s);
    out.throttled = { frames, maxInFlight, reads: got, step: s.stepCount };
    s.destroy();
  }
  // 4. calls while a sample is in flight, then destroy with a sample still pending
  {
    const a = new SolverGPU(dims, params);
    const b = new SolverGPU(dims, params);
    a.uploadFlags(flags);
    b.uploadFlags(flags);
    const p = a.stepAndSample(200);  // not awaited
    a.encodeSteps(30);               // after the sample's steps, in call order
    const q = a.stepAndSample(0);
    const r1 = await p;
    const r2 = await q;
    const rb1 = await b.stepAndSample(200);
    b.encodeSteps(30);
    const rb2 = await b.stepAndSample(0);
    out.race = { a1: [r1.step, r1.fx], a2: [r2.step, r2.fx], b1: [rb1.step, rb1.fx], b2: [rb2.step, rb2.fx] };
    const c = new SolverGPU(dims, params);
    c.uploadFlags(flags);
    const pc = c.stepAndSample(300);
    c.destroy();  // deferred until the sample completes
    let threw = false;
    try { c.encodeSteps(1); } catch (e) { threw = true; }
    const rc = await pc;
    out.destroyed = { step: rc.step, unstable: rc.unstable, threwAfterDestroy: threw };
    a.destroy();
    b.destroy();
  }
  console.log(JSON.stringify(out));
}

main().catch((e) => { console.error(e); process.exit(1); });
