// CPU restatement of the mesh voxelizer (TEST INFRASTRUCTURE ONLY; see lbm_oracle.cpp header).
//
// Follows the reference's GPU voxelizer, src/voxelize/wgslVoxel.ts:33-322 with the host sequence of
// src/voxelize/VoxelizerGPU.ts:48-151 (crossings -> scans -> shell -> finalize -> distances ->
// wheel tags -> projection), in FP32 with the same operation order as the HIP kernels and no
// contraction, so integer outputs (flags, frontal, counts, bbox) can be compared exactly.
// The reference's CPU twin (src/voxelize/voxelizeCPU.ts:32-199) is the same algorithm in FP64
// without wheel tags.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

namespace {

struct F3 { float x, y, z; };
inline F3 sub3(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float dot3(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline float len3(F3 a) { return std::sqrt(dot3(a, a)); }
inline F3 cross3(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float comp(F3 p, int a) { return a == 0 ? p.x : (a == 1 ? p.y : p.z); }

// Ericson, Real-Time Collision Detection 5.1.5 (voxelizeCPU.ts:11-30, wgslVoxel.ts:211-227)
F3 closest_pt(F3 p, F3 a, F3 b, F3 c) {
  F3 ab = sub3(b, a), ac = sub3(c, a), ap = sub3(p, a);
  float d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return a;
  F3 bp = sub3(p, b);
  float d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return b;
  float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    float v = d1 / (d1 - d3);
    return {a.x + ab.x * v, a.y + ab.y * v, a.z + ab.z * v};
  }
  F3 cp = sub3(p, c);
  float d5 = dot3(ab, cp), d6 = dot3(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return c;
  float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    float w = d2 / (d2 - d6);
    return {a.x + ac.x * w, a.y + ac.y * w, a.z + ac.z * w};
  }
  float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return {b.x + (c.x - b.x) * w, b.y + (c.y - b.y) * w, b.z + (c.z - b.z) * w};
  }
  float den = 1.0f / ((va + vb) + vc);
  float v = vb * den, w = vc * den;
  return {(a.x + ab.x * v) + ac.x * w, (a.y + ab.y * v) + ac.y * w, (a.z + ab.z * v) + ac.z * w};
}

inline int popc3(uint32_t v) { return (int)(v & 1u) + (int)((v >> 1) & 1u) + (int)((v >> 2) & 1u); }

}  // namespace

extern "C" {

// positions: 3 floats per vertex; indices: 3 per triangle; M: column-major 4x4; wheels: 8 floats per
// wheel (cx, cy, cz, ax, ay, az, r, hw). Outputs: flags[n], sdf[n] (encoded ~bits, 0 = unknown),
// info[8] = frontal, solid, minx, miny, minz, maxx, maxy, maxz (max inclusive).
void or_voxelize_mesh(const float* positions, int nv, const uint32_t* indices, int nt, const float* M, const float* wheels,
                      int nwheels, int nx, int ny, int nz, uint32_t* flags, uint32_t* sdf, int64_t* info) {
  (void)nv;
  const int64_t n = (int64_t)nx * ny * nz;
  const int dims[3] = {nx, ny, nz};
  auto cell = [&](int x, int y, int z) { return (int64_t)x + (int64_t)nx * ((int64_t)y + (int64_t)ny * z); };
  auto vtx = [&](int64_t i) {
    const uint32_t k = indices[i];
    const float x = positions[3 * k], y = positions[3 * k + 1], z = positions[3 * k + 2];
    return F3{((M[0] * x + M[4] * y) + M[8] * z) + M[12], ((M[1] * x + M[5] * y) + M[9] * z) + M[13],
              ((M[2] * x + M[6] * y) + M[10] * z) + M[14]};
  };
  std::memset(sdf, 0, 4 * n);
  if (nt == 0) {
    std::memset(flags, 0, 4 * n);
    for (int k = 0; k < 8; k++) info[k] = 0;
    return;
  }
  std::vector<int> hits(3 * (size_t)n, 0);
  std::vector<uint32_t> vote((size_t)n, 0u);
  // 1. crossings
  for (int t = 0; t < nt; t++) {
    F3 p0 = vtx(3 * (int64_t)t), p1 = vtx(3 * (int64_t)t + 1), p2 = vtx(3 * (int64_t)t + 2);
    for (int a = 0; a < 3; a++) {
      const int ua = (a + 1) % 3, va = (a + 2) % 3;
      const float Ax = comp(p0, ua), Ay = comp(p0, va), Bx = comp(p1, ua), By = comp(p1, va), Cx = comp(p2, ua), Cy = comp(p2, va);
      const float area = (Bx - Ax) * (Cy - Ay) - (By - Ay) * (Cx - Ax);
      if (std::fabs(area) < 1e-12f) continue;
      const int sgn = area > 0.0f ? -1 : 1;
      const int nu = dims[ua], nvv = dims[va], na = dims[a];
      const float lox = std::fmin(std::fmin(Ax, Bx), Cx), hix = std::fmax(std::fmax(Ax, Bx), Cx);
      const float loy = std::fmin(std::fmin(Ay, By), Cy), hiy = std::fmax(std::fmax(Ay, By), Cy);
      const int u0 = std::max((int)std::ceil(lox - 0.5f), 0), u1 = std::min((int)std::floor(hix - 0.5f), nu - 1);
      const int v0 = std::max((int)std::ceil(loy - 0.5f), 0), v1 = std::min((int)std::floor(hiy - 0.5f), nvv - 1);
      const float ha = comp(p0, a), hb = comp(p1, a), hc = comp(p2, a);
      const float inv = 1.0f / area;
      for (int iu = u0; iu <= u1; iu++)
        for (int iv = v0; iv <= v1; iv++) {
          const float Px = ((float)iu + 0.5f) + 1.37e-4f, Py = ((float)iv + 0.5f) + 2.71e-4f;
          const float w0 = ((Bx - Px) * (Cy - Py) - (By - Py) * (Cx - Px)) * inv;
          const float w1 = ((Cx - Px) * (Ay - Py) - (Cy - Py) * (Ax - Px)) * inv;
          const float w2 = (1.0f - w0) - w1;
          if (w0 < 0.0f || w1 < 0.0f || w2 < 0.0f) continue;
          const float h = (w0 * ha + w1 * hb) + w2 * hc;
          int k = (int)std::floor(h - 0.5f) + 1;
          if (k >= na) continue;
          k = std::max(k, 0);
          const int64_t c = a == 0 ? cell(k, iu, iv) : (a == 1 ? cell(iv, k, iu) : cell(iu, iv, k));
          hits[(size_t)a * n + c] += sgn;
        }
    }
  }
  // 2. scans
  for (int a = 0; a < 3; a++) {
    const int ua = (a + 1) % 3, va = (a + 2) % 3;
    for (int iv = 0; iv < dims[va]; iv++)
      for (int iu = 0; iu < dims[ua]; iu++) {
        int s = 0;
        for (int k = 0; k < dims[a]; k++) {
          const int64_t c = a == 0 ? cell(k, iu, iv) : (a == 1 ? cell(iv, k, iu) : cell(iu, iv, k));
          s += hits[(size_t)a * n + c];
          if (s != 0) vote[c] |= 1u << a;
        }
      }
  }
  // 3. shell
  const bool is2D = nz == 1;
  for (int t = 0; t < nt; t++) {
    F3 p0 = vtx(3 * (int64_t)t);
    F3 e1 = sub3(vtx(3 * (int64_t)t + 1), p0), e2 = sub3(vtx(3 * (int64_t)t + 2), p0);
    const float m = std::fmax(std::fmax(len3(e1), len3(e2)), len3(sub3(e2, e1)));
    const int steps = (int)std::fmin(std::fmax(std::ceil(m / 0.4f), 1.0f), 384.0f);
    const float fs = (float)steps;
    const F3 cr = cross3(e1, e2);
    const float cl = len3(cr);
    if (cl < 1e-12f) continue;
    const F3 nrm{cr.x / cl, cr.y / cl, cr.z / cl};
    const float tol = 0.5f * std::fmax(std::fmax(std::fabs(nrm.x), std::fabs(nrm.y)), std::fabs(nrm.z));
    for (int i = 0; i <= steps; i++)
      for (int j = 0; j <= steps - i; j++) {
        const float si = (float)i / fs, sj = (float)j / fs;
        const F3 p{(p0.x + e1.x * si) + e2.x * sj, (p0.y + e1.y * si) + e2.y * sj, (p0.z + e1.z * si) + e2.z * sj};
        if (p.x < 0.0f || p.y < 0.0f || p.z < 0.0f || p.x >= (float)nx || p.y >= (float)ny || p.z >= (float)nz) continue;
        const float cx = std::floor(p.x), cy = std::floor(p.y), cz = std::floor(p.z);
        F3 ctr{cx + 0.5f, cy + 0.5f, cz + 0.5f};
        float tt = tol;
        if (is2D) {
          const float n2 = std::sqrt(nrm.x * nrm.x + nrm.y * nrm.y);
          if (n2 < 1e-4f) continue;
          ctr.z = p.z;
          tt = 0.5f * std::fmax(std::fabs(nrm.x), std::fabs(nrm.y));
          if (std::fabs(p.z - 0.5f) > 0.5f) continue;
        }
        if (std::fabs(dot3(sub3(ctr, p0), nrm)) > tt) continue;
        vote[cell((int)cx, (int)cy, (int)cz)] |= 8u;
      }
  }
  // 4. finalize
  auto votes = [&](uint32_t v) { return popc3(v) >= 2; };
  for (int64_t id = 0; id < n; id++) {
    const uint32_t v = vote[id];
    bool solid = votes(v);
    if (!solid && (v & 8u)) {
      const int x = (int)(id % nx), y = (int)((id / nx) % ny), z = (int)(id / ((int64_t)nx * ny));
      bool near = false;
      if (x > 0) near = near || votes(vote[id - 1]);
      if (x + 1 < nx) near = near || votes(vote[id + 1]);
      if (y > 0) near = near || votes(vote[id - nx]);
      if (y + 1 < ny) near = near || votes(vote[id + nx]);
      if (z > 0) near = near || votes(vote[id - (int64_t)nx * ny]);
      if (z + 1 < nz) near = near || votes(vote[id + (int64_t)nx * ny]);
      solid = !near;
    }
    flags[id] = solid ? 1u : 0u;
  }
  // 5. distances
  for (int t = 0; t < nt; t++) {
    const F3 a = vtx(3 * (int64_t)t), b = vtx(3 * (int64_t)t + 1), c = vtx(3 * (int64_t)t + 2);
    const F3 e1 = sub3(b, a), e2 = sub3(c, a);
    if (len3(cross3(e1, e2)) < 1e-12f) continue;
    const float m = std::fmax(std::fmax(len3(e1), len3(e2)), len3(sub3(e2, e1)));
    const int steps = (int)std::fmin(std::fmax(std::ceil(m / 0.7f), 1.0f), 256.0f);
    const float fs = (float)steps;
    const int R = 2;
    for (int i = 0; i <= steps; i++)
      for (int j = 0; j <= steps - i; j++) {
        const float si = (float)i / fs, sj = (float)j / fs;
        const F3 p{(a.x + e1.x * si) + e2.x * sj, (a.y + e1.y * si) + e2.y * sj, (a.z + e1.z * si) + e2.z * sj};
        if (is2D && std::fabs(p.z - 0.5f) > 2.5f) continue;
        const int bx = (int)std::floor(p.x), by = (int)std::floor(p.y), bz = (int)std::floor(p.z);
        for (int dz = -R; dz <= R; dz++) {
          if (is2D && dz != 0) continue;
          for (int dy = -R; dy <= R; dy++)
            for (int dx = -R; dx <= R; dx++) {
              const int qx = bx + dx, qy = by + dy, qz = is2D ? 0 : bz + dz;
              if (qx < 0 || qy < 0 || qz < 0 || qx >= nx || qy >= ny || qz >= nz) continue;
              const F3 ctr{(float)qx + 0.5f, (float)qy + 0.5f, (float)qz + 0.5f};
              const float d = len3(sub3(ctr, closest_pt(ctr, a, b, c)));
              if (d > 2.5f) continue;
              uint32_t bits;
              std::memcpy(&bits, &d, 4);
              const int64_t id = cell(qx, qy, qz);
              sdf[id] = std::max(sdf[id], ~bits);
            }
        }
      }
  }
  // 6. wheel tags
  for (int64_t id = 0; id < n && nwheels > 0; id++) {
    if (flags[id] == 0u) continue;
    const F3 p{(float)(id % nx) + 0.5f, (float)((id / nx) % ny) + 0.5f, (float)(id / ((int64_t)nx * ny)) + 0.5f};
    for (int k = 0; k < std::min(nwheels, 16); k++) {
      const float* w = wheels + 8 * k;
      const float inv_r = 1.0f / std::max(w[6], 1e-3f);
      const F3 d{p.x - w[0], p.y - w[1], p.z - w[2]};
      const float al = dot3(d, F3{w[3], w[4], w[5]});
      const float rad = len3(F3{d.x - al * w[3], d.y - al * w[4], d.z - al * w[5]});
      if (std::fabs(al) <= w[7] + 0.6f && rad <= 1.0f / inv_r + 0.75f) {
        flags[id] = 1u | ((uint32_t)(k + 1) << 8);
        break;
      }
    }
  }
  // 7. projection (wgslVoxel.ts:295-322; max kept inclusive here)
  int64_t frontal = 0, solid = 0;
  int64_t mn[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, mx[3] = {0, 0, 0};
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++) {
      bool hit = false;
      for (int x = 0; x < nx; x++)
        if (flags[cell(x, y, z)] != 0u) {
          if (!hit) mn[0] = std::min<int64_t>(mn[0], x);
          mx[0] = std::max<int64_t>(mx[0], x);
          hit = true;
          solid++;
        }
      if (hit) {
        frontal++;
        mn[1] = std::min<int64_t>(mn[1], y); mx[1] = std::max<int64_t>(mx[1], y);
        mn[2] = std::min<int64_t>(mn[2], z); mx[2] = std::max<int64_t>(mx[2], z);
      }
    }
  info[0] = frontal;
  info[1] = solid;
  for (int a = 0; a < 3; a++) { info[2 + a] = mn[a]; info[5 + a] = mx[a]; }
}

}  // extern "C"
