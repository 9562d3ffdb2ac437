// Headless validation harness for the Node host, with the reference's case ids and pass bands
// (src/test/harness.ts:200-216, scripts/harness.mjs:11-54): runs the selected cases on the MI355X
// solver and exits non-zero when one fails.
//   node windtunnel_vis_amd/node/harness.mjs --case=freestream,cylinder --valscale=0.5 [--prec=f16]
import { SolverGPU } from './solver.mjs';
import { makeCases } from './validation.mjs';

const args = Object.fromEntries(process.argv.slice(2).map((a) => {
  const m = a.match(/^--([^=]+)=?(.*)$/);
  return m ? [m[1], m[2] === '' ? true : m[2]] : [a, true];
}));
const which = String(args.case || 'freestream,cylinder').split(',');
const scale = parseFloat(args.valscale || '1');
const prec = args.prec || 'f32';

async function main() {
  const t0 = Date.now();
  const results = [];
  for (const c of makeCases(scale)) {
    if (!which.includes(c.id)) continue;
    console.log(`running ${c.name}`);
    try {
      const r = await c.run((d, p) => new SolverGPU(d, { ...p, precision: prec }));
      console.log(JSON.stringify(r));
      results.push(r);
    } catch (e) {
      console.log(`ERROR ${e}`);
      results.push({ id: c.id, passed: false, error: String(e) });
    }
  }
  const failed = results.filter((r) => r.passed === false);
  console.log(`finished in ${((Date.now() - t0) / 1000).toFixed(1)} s: ${results.length - failed.length}/${results.length} passed`);
  return failed.length ? 1 : 0;
}

// Node 12 has no top-level await
main().then((code) => process.exit(code), (e) => { console.error(e); process.exit(2); });
