// Writes js_number_vectors.json: JavaScript's own Number formatting on a fixed set of doubles, the
// fixture for the CSV / VTK writers' number printing (String(v), +v.toPrecision(7), and the
// big-endian float32 that DataView.setFloat32 stores). Run: node make_js_number_vectors.js
'use strict';
const fs = require('fs');
const path = require('path');

const vals = [0, -0, 1, -1, 0.5, 2.5, 0.1, 0.2 + 0.1, 1 / 3, 2 / 3, Math.PI, -Math.E, 100, 1e6, 1234567, 1234567.5,
  1234568.5, -1234567.5, 12345675, 0.12345675, 1.0000005, 9.9999995, 99999995, 1e21, 1e20, 9.99e20, 1.5e21, 123456789012345680000,
  1e-6, 1e-7, 9.99e-7, 1.5e-7, 0.000001234, 0.0000012345678, 1e-300, 5e-324, 1.7976931348623157e308, 4.35e-5, 0.00015,
  6.02214076e23, 299792458, 0.06, 0.0006, 4.5976628609671994, 5.230868081036491, 3.0000000000000004, 255.99999999999997];
let s = 0x9e3779b9 >>> 0;
function rnd() {  // xorshift32
  s ^= s << 13; s >>>= 0; s ^= s >>> 17; s ^= s << 5; s >>>= 0;
  return s / 4294967296;
}
for (let k = 0; k < 400; k++) {
  const e = Math.floor(rnd() * 60) - 30;
  const m = (rnd() * 9 + 1) * (rnd() < 0.3 ? -1 : 1);
  vals.push(m * Math.pow(10, e));
}
for (let k = 0; k < 100; k++) {  // decimal ties at the 7th digit
  const d = Math.floor(rnd() * 9000000) + 1000000;
  const e = Math.floor(rnd() * 10) - 5;
  vals.push((d + 0.5) * Math.pow(10, e));
}
const dv = new DataView(new ArrayBuffer(8));
const out = vals.map((v) => {
  dv.setFloat64(0, v, false);
  const bits = dv.getBigUint64(0, false).toString(16).padStart(16, '0');
  dv.setFloat32(0, v, false);
  const f32 = dv.getUint32(0, false).toString(16).padStart(8, '0');
  return { bits, str: String(v), p7: String(+v.toPrecision(7)), f32be: f32 };
});
fs.writeFileSync(path.join(__dirname, 'js_number_vectors.json'), JSON.stringify(out, null, 0) + '\n');
console.log(`${out.length} vectors`);
