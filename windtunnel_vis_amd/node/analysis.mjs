// Host-side measurement for the Node host: aerodynamic coefficients, batch-means statistics and
// the mean-crossing shedding frequency, restating src/analysis/aero.ts:4-52 and
// src/analysis/strouhal.ts:5-40.

/** C_D / C_L / C_S from a lattice force (aero.ts:4-7). In 2D `area` is the frontal height in cells. */
export function coefficients(fx, fy, fz, U, area, rho = 1) {
  const q = 0.5 * rho * U * U * Math.max(area, 1e-9);
  return { cd: fx / q, cl: fy / q, cs: fz / q };
}

/** Mean and standard error over the last `window` time units, 8 batch means (aero.ts:33-52). */
export function windowStats(t, v, window) {
  const n = t.length;
  if (n < 16) return { mean: NaN, se: NaN, span: 0, n };
  const tEnd = t[n - 1];
  let i0 = n - 1;
  while (i0 > 0 && t[i0 - 1] >= tEnd - window) i0--;
  const span = tEnd - t[i0];
  const B = 8;
  const sums = new Array(B).fill(0), cnt = new Array(B).fill(0);
  let tot = 0, totN = 0;
  for (let i = i0; i < n; i++) {
    const b = Math.min(B - 1, Math.floor(((t[i] - t[i0]) / Math.max(span, 1e-9)) * B));
    sums[b] += v[i]; cnt[b]++; tot += v[i]; totN++;
  }
  const mean = tot / totN;
  const bm = sums.map((s, k) => (cnt[k] ? s / cnt[k] : NaN)).filter((x) => isFinite(x));
  if (bm.length < 4) return { mean, se: NaN, span, n: totN };
  const varB = bm.reduce((a, x) => a + (x - mean) ** 2, 0) / (bm.length - 1);
  return { mean, se: Math.sqrt(varB / bm.length), span, n: totN };
}

/** Dominant frequency from upward mean crossings with 15 % hysteresis (strouhal.ts:5-30). */
export function crossingFrequency(t, v) {
  const n = v.length;
  if (n < 8) return { f: NaN, cycles: 0, amplitude: NaN };
  let mean = 0;
  for (const x of v) mean += x;
  mean /= n;
  let lo = Infinity, hi = -Infinity;
  for (const x of v) { lo = Math.min(lo, x); hi = Math.max(hi, x); }
  const amp = (hi - lo) / 2;
  const hyst = amp * 0.15;
  const ups = [];
  let armed = v[0] < mean - hyst;
  for (let i = 1; i < n; i++) {
    if (v[i] < mean - hyst) armed = true;
    if (armed && v[i - 1] < mean && v[i] >= mean) {
      const a = (mean - v[i - 1]) / (v[i] - v[i - 1]);
      ups.push(t[i - 1] + a * (t[i] - t[i - 1]));
      armed = false;
    }
  }
  if (ups.length < 4) return { f: NaN, cycles: Math.max(0, ups.length - 1), amplitude: amp };
  const cycles = ups.length - 1;
  return { f: cycles / (ups[ups.length - 1] - ups[0]), cycles, amplitude: amp };
}

/** Schiller-Naumann sphere drag (strouhal.ts:33-35). */
export const sphereCdRef = (re) => (24 / re) * (1 + 0.15 * Math.pow(re, 0.687));
/** Williamson cylinder Strouhal fit (strouhal.ts:38-40). */
export const cylinderStRef = (re) => 0.2175 - 5.106 / re;
