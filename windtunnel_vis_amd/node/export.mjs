// Result and field writers of the Node host (src/analysis/export.ts), returning the bytes instead
// of starting a browser download: csv() a string, vtk() a Buffer. Same layouts as the Python
// writers in windtunnel_vis_amd/export.py.

/** '# ' comments, header, rows; numbers at 7 significant digits, non-finite as empty cells. */
export function csv(header, rows, comments = []) {
  const lines = comments.map((c) => `# ${c}`);
  lines.push(header.join(','));
  for (const r of rows) {
    lines.push(r.map((v) => {
      if (typeof v !== 'number') return v;
      return Number.isFinite(v) ? String(Number(v.toPrecision(7))) : '';
    }).join(','));
  }
  return lines.join('\n') + '\n';
}

/** Export stride keeping at most ~maxPoints points (backend/webgpu.ts:133-137). */
export function strideFor(n, maxPoints) {
  return Math.max(1, Math.ceil(Math.cbrt(n / maxPoints)));
}

/**
 * Legacy VTK (binary, big-endian) STRUCTURED_POINTS file of a FieldExport: velocities / U,
 * vorticity L / U, Cp of the instantaneous and mean density, Q L^2 / U^2 (export.ts:29-57).
 */
export function vtk(f, o) {
  const [nx, ny, nz] = f.dims;
  const n = nx * ny * nz;
  const h = o.dx * f.stride;
  const chunks = [];
  const text = (s) => chunks.push(Buffer.from(s, 'utf8'));
  const block = (comps, value) => {
    const b = Buffer.alloc(4 * n * comps);
    for (let i = 0; i < n; i++) for (let k = 0; k < comps; k++) b.writeFloatBE(value(i, k), 4 * (i * comps + k));
    chunks.push(b);
    text('\n');
  };
  const cp = (rho) => (2 * (rho - o.rhoRef)) / (3 * o.U * o.U);
  const half = 0.5 * o.dx;
  text(`# vtk DataFile Version 3.0\n${o.title.slice(0, 250)}\nBINARY\nDATASET STRUCTURED_POINTS\n`);
  text(`DIMENSIONS ${nx} ${ny} ${nz}\nORIGIN ${half} ${half} ${nz > 1 ? half : 0}\nSPACING ${h} ${h} ${h}\nPOINT_DATA ${n}\n`);
  text('VECTORS velocity float\n');
  block(3, (i, k) => f.vel[4 * i + k] / o.U);
  text('VECTORS mean_velocity float\n');
  block(3, (i, k) => f.mean[4 * i + k] / o.U);
  text('VECTORS vorticity float\n');
  block(3, (i, k) => (f.vort[4 * i + k] * o.L) / o.U);
  text('SCALARS Cp float 1\nLOOKUP_TABLE default\n');
  block(1, (i) => cp(f.vel[4 * i + 3]));
  text('SCALARS mean_Cp float 1\nLOOKUP_TABLE default\n');
  block(1, (i) => cp(f.mean[4 * i + 3]));
  text('SCALARS Q_criterion float 1\nLOOKUP_TABLE default\n');
  block(1, (i) => (f.vort[4 * i + 3] * o.L * o.L) / (o.U * o.U));
  return Buffer.concat(chunks);
}
