// Validation cases with the reference's setups and pass bands (src/analysis/validation.ts:60-184),
// run on any SolverFactory (here the MI355X solver).
import { coefficients, crossingFrequency, cylinderStRef, sphereCdRef } from './analysis.mjs';

/** Analytic circle / sphere flags and wall distances (validation.ts:35-53). */
export function roundFlags(d, c, r, cylinder) {
  const f = new Uint32Array(d.nx * d.ny * d.nz);
  const sdf = new Float32Array(d.nx * d.ny * d.nz).fill(-1);
  const proj = new Uint8Array(d.ny * d.nz);
  for (let z = 0; z < d.nz; z++)
    for (let y = 0; y < d.ny; y++)
      for (let x = 0; x < d.nx; x++) {
        const dx = x + 0.5 - c[0], dy = y + 0.5 - c[1], dz = cylinder ? 0 : z + 0.5 - c[2];
        const dist = Math.abs(Math.sqrt(dx * dx + dy * dy + dz * dz) - r);
        if (dist < 2.5) sdf[x + d.nx * (y + d.ny * z)] = dist;
        if (dx * dx + dy * dy + dz * dz < r * r) {
          f[x + d.nx * (y + d.ny * z)] = 1;
          proj[y + d.ny * z] = 1;
        }
      }
  let frontal = 0;
  for (const p of proj) frontal += p;
  return { flags: f, frontal, sdf };
}

export function makeCases(scale = 1) {
  return [
    {
      id: 'freestream',
      name: 'Free-stream uniformity',
      async run(make) {
        const t0 = Date.now();
        const d = { nx: 256, ny: 96, nz: 1 };
        const U = 0.08;
        const s = make(d, { u: U, nu: 0.01, cs: 0.1, ground: 'freeslip', sides: 'freeslip', spongeNu: 0.05, spongeStart: 0.85, emaAlpha: 0.02 });
        s.uploadFlags(new Uint32Array(d.nx * d.ny));
        let st = await s.stepAndSample(1);
        for (let i = 0; i < 20; i++) st = await s.stepAndSample(50);
        s.destroy();
        const err = Math.abs(st.maxU - U) / U;
        const drho = Math.abs(st.rhoRef - 1);
        return {
          id: 'freestream', passed: err < 0.01 && drho < 0.01 && !st.unstable, seconds: (Date.now() - t0) / 1000,
          metrics: [{ label: 'max |u| / U', value: st.maxU / U }, { label: 'rho at inlet plane', value: st.rhoRef }],
        };
      },
    },
    {
      id: 'cylinder',
      name: 'Cylinder, Re = 100',
      async run(make) {
        const t0 = Date.now();
        const D = Math.round(24 * scale);
        const d = { nx: Math.round(25 * D), ny: 10 * D, nz: 1 };
        const U = 0.06, Re = 100, nu = (U * D) / Re;
        const s = make(d, { u: U, nu, cs: 0, ground: 'freeslip', sides: 'freeslip', spongeNu: 0.03, spongeStart: 0.85, emaAlpha: 0.0 });
        const { flags, sdf } = roundFlags(d, [6 * D, d.ny / 2 + 0.37, 0], D / 2, true);
        s.uploadFlags(flags);
        s.uploadSdf(sdf);
        const period = D / (0.17 * U);
        const total = Math.round(period * 16);
        const measureFrom = Math.round(period * 6);
        const chunk = Math.max(5, Math.round(period / 40));
        const tt = [], cl = [], cd = [];
        for (let step = 0; step < total; step += chunk) {
          const st = await s.stepAndSample(chunk);
          if (st.unstable) break;
          const c = coefficients(st.fx, st.fy, 0, U, D);
          if (step >= measureFrom) { tt.push(st.step); cl.push(c.cl); cd.push(c.cd); }
        }
        s.destroy();
        const { f, cycles, amplitude } = crossingFrequency(tt, cl);
        const St = (f * D) / U;
        const cdMean = cd.reduce((a, b) => a + b, 0) / Math.max(cd.length, 1);
        return {
          id: 'cylinder', passed: St >= 0.16 && St <= 0.2 && cycles >= 4, seconds: (Date.now() - t0) / 1000,
          metrics: [{ label: 'St', value: St, expected: `0.16-0.20 (Williamson ${cylinderStRef(Re).toFixed(3)})` },
            { label: 'mean C_D', value: cdMean }, { label: 'C_L amplitude', value: amplitude }, { label: 'cycles', value: cycles }],
        };
      },
    },
    {
      id: 'sphere',
      name: 'Sphere, Re = 100',
      async run(make) {
        const t0 = Date.now();
        const D = Math.round(20 * scale);
        const d = { nx: Math.round(9.6 * D), ny: Math.round(4.8 * D), nz: Math.round(4.8 * D) };
        const U = 0.05, Re = 100, nu = (U * D) / Re;
        const s = make(d, { u: U, nu, cs: 0, ground: 'freeslip', sides: 'freeslip', spongeNu: 0.03, spongeStart: 0.85, emaAlpha: 0 });
        const { flags, frontal, sdf } = roundFlags(d, [2.8 * D, d.ny / 2 + 0.37, d.nz / 2 + 0.21], D / 2, false);
        s.uploadFlags(flags);
        s.uploadSdf(sdf);
        const flowThrough = d.nx / U;
        const total = Math.round(flowThrough * 2.2);
        const measureFrom = Math.round(flowThrough * 1.4);
        const tt = [], cds = [];
        for (let step = 0; step < total; step += 50) {
          const st = await s.stepAndSample(50);
          if (st.unstable) break;
          tt.push(st.step);
          cds.push(coefficients(st.fx, st.fy, st.fz, U, frontal).cd);
        }
        s.destroy();
        const sel = cds.filter((_, i) => tt[i] >= measureFrom);
        const cd = sel.reduce((a, b) => a + b, 0) / Math.max(sel.length, 1);
        const ref = sphereCdRef(Re);
        const err = (cd - ref) / ref;
        return {
          id: 'sphere', passed: Math.abs(err) < 0.2, seconds: (Date.now() - t0) / 1000,
          metrics: [{ label: 'mean C_D', value: cd, expected: `${ref.toFixed(3)} +- 20 %` }, { label: 'deviation', value: err }],
        };
      },
    },
  ];
}
