// Voxelizers: analytic signed-distance bodies and (see below) triangle meshes, producing the solid
// flags, wall distances and wheel tags the LBM kernels consume, plus the frontal projection and
// bounding box (VoxelInfo).
#include <vector>
#include "voxel.hpp"

namespace wt {

namespace {

constexpr int VB = 256;

struct PrimDev {
  int kind, pad;
  float c[3];
  float r[9];
  float h[3];
  float radius, param;
};

__device__ inline float len3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

// Signed distance of one primitive at grid point p (negative inside).
__device__ float prim_sdf(const PrimDev& P, float px, float py, float pz) {
  const float dx = px - P.c[0], dy = py - P.c[1], dz = pz - P.c[2];
  // local = R^T d (R row-major: R[row*3+col])
  const float lx = P.r[0] * dx + P.r[3] * dy + P.r[6] * dz;
  const float ly = P.r[1] * dx + P.r[4] * dy + P.r[7] * dz;
  const float lz = P.r[2] * dx + P.r[5] * dy + P.r[8] * dz;
  if (P.kind == WT_PRIM_SPHERE) return len3(lx, ly, lz) - P.h[0];
  if (P.kind == WT_PRIM_CYLINDER) {  // axis = local z, radius h[0], half width h[1]
    const float a = sqrtf(lx * lx + ly * ly) - P.h[0], b = fabsf(lz) - P.h[1];
    return fminf(fmaxf(a, b), 0.0f) + sqrtf(fmaxf(a, 0.0f) * fmaxf(a, 0.0f) + fmaxf(b, 0.0f) * fmaxf(b, 0.0f));
  }
  // rounded box (and wedge base): half extents h include the rounding radius
  const float r = fminf(P.radius, fminf(P.h[0], fminf(P.h[1], P.h[2])));
  const float qx = fabsf(lx) - (P.h[0] - r), qy = fabsf(ly) - (P.h[1] - r), qz = fabsf(lz) - (P.h[2] - r);
  float d = len3(fmaxf(qx, 0.f), fmaxf(qy, 0.f), fmaxf(qz, 0.f)) + fminf(fmaxf(qx, fmaxf(qy, qz)), 0.0f) - r;
  if (P.kind == WT_PRIM_WEDGE) {
    // top face slopes down toward -x (the nose): height at local x is h.y - param * (h.x - x)
    const float s = P.param;
    const float nl = sqrtf(1.0f + s * s);
    const float plane = (ly - (P.h[1] - s * (P.h[0] - lx))) / nl;
    d = fmaxf(d, plane);
  }
  return d;
}

__global__ __launch_bounds__(VB) void k_vox_sdf(const PrimDev* __restrict__ prims, int count, uint32_t* __restrict__ flags,
                                                uint32_t* __restrict__ sdf, int nx, int ny, int nz, int64_t n, KWheels W) {
  const int64_t id = (int64_t)blockIdx.x * VB + threadIdx.x;
  if (id >= n) return;
  const int x = (int)(id % nx), y = (int)((id / nx) % ny), z = (int)(id / ((int64_t)nx * ny));
  const float px = x + 0.5f, py = y + 0.5f, pz = z + 0.5f;
  float d = 1e30f;
  for (int k = 0; k < count; k++) d = fminf(d, prim_sdf(prims[k], px, py, pz));
  uint32_t fl = d < 0.0f ? 1u : 0u;
  if (fl) {  // wheel tag, wgslVoxel.ts:268-290 (first wheel whose tagging cylinder holds the centre)
    for (int k = 0; k < W.count; k++) {
      const float4 a = W.w[2 * k], b = W.w[2 * k + 1];
      const float ddx = px - a.x, ddy = py - a.y, ddz = pz - a.z;
      const float al = ddx * b.x + ddy * b.y + ddz * b.z;
      const float rad = len3(ddx - al * b.x, ddy - al * b.y, ddz - al * b.z);
      if (fabsf(al) <= b.w + 0.6f && rad <= 1.0f / a.w + 0.75f) { fl = 1u | ((uint32_t)(k + 1) << 8); break; }
    }
  }
  flags[id] = fl;
  const float ad = fabsf(d);
  sdf[id] = ad < 2.5f ? ~__float_as_uint(ad) : 0u;
}

// Frontal projection + bbox (wgslVoxel.ts:295-322): one thread per (y, z) column.
__global__ __launch_bounds__(VB) void k_project(const uint32_t* __restrict__ flags, int nx, int ny, int nz,
                                                unsigned long long* __restrict__ info) {
  const int64_t col = (int64_t)blockIdx.x * VB + threadIdx.x;
  if (col >= (int64_t)ny * nz) return;
  const int y = (int)(col % ny), z = (int)(col / ny);
  bool hit = false;
  unsigned long long cnt = 0;
  int xmin = 0, xmax = 0;
  for (int x = 0; x < nx; x++) {
    if (flags[(int64_t)x + (int64_t)nx * ((int64_t)y + (int64_t)ny * z)] != 0u) {
      if (!hit) xmin = x;
      xmax = x;
      hit = true;
      cnt++;
    }
  }
  if (hit) {
    atomicAdd(&info[0], 1ull);
    atomicAdd(&info[1], cnt);
    atomicMin(&info[2], (unsigned long long)xmin); atomicMax(&info[5], (unsigned long long)xmax);
    atomicMin(&info[3], (unsigned long long)y); atomicMax(&info[6], (unsigned long long)y);
    atomicMin(&info[4], (unsigned long long)z); atomicMax(&info[7], (unsigned long long)z);
  }
}

}  // namespace

int project_info(const VoxCtx& v, int64_t* info8) {
  unsigned long long h[8] = {0, 0, ~0ull, ~0ull, ~0ull, 0, 0, 0};
  unsigned long long* d = nullptr;
  if (hipMallocAsync((void**)&d, sizeof(h), v.stream) != hipSuccess) return -1;
  if (hipMemcpyAsync(d, h, sizeof(h), hipMemcpyHostToDevice, v.stream) != hipSuccess) return -1;
  const int64_t cols = (int64_t)v.ny * v.nz;
  hipLaunchKernelGGL(k_project, dim3((unsigned)((cols + VB - 1) / VB)), dim3(VB), 0, v.stream, v.flags, v.nx, v.ny, v.nz, d);
  if (hipGetLastError() != hipSuccess) return -1;
  if (hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, v.stream) != hipSuccess) return -1;
  if (hipFreeAsync(d, v.stream) != hipSuccess) return -1;
  if (hipStreamSynchronize(v.stream) != hipSuccess) return -1;
  for (int k = 0; k < 8; k++) info8[k] = (int64_t)h[k];
  return 0;
}

int voxelize_sdf(const VoxCtx& v, const wt_prim* prims, int count, int64_t* info8) {
  std::vector<PrimDev> hp((size_t)std::max(count, 1));
  for (int k = 0; k < count; k++) {
    PrimDev& d = hp[k];
    d.kind = prims[k].kind;
    d.pad = 0;
    for (int a = 0; a < 3; a++) { d.c[a] = prims[k].center[a]; d.h[a] = prims[k].half[a]; }
    for (int a = 0; a < 9; a++) d.r[a] = prims[k].rot[a];
    d.radius = prims[k].radius;
    d.param = prims[k].param;
  }
  PrimDev* dp = nullptr;
  if (hipMallocAsync((void**)&dp, sizeof(PrimDev) * hp.size(), v.stream) != hipSuccess) return -1;
  if (hipMemcpyAsync(dp, hp.data(), sizeof(PrimDev) * hp.size(), hipMemcpyHostToDevice, v.stream) != hipSuccess) return -1;
  hipLaunchKernelGGL(k_vox_sdf, dim3((unsigned)((v.n + VB - 1) / VB)), dim3(VB), 0, v.stream, dp, count, v.flags, v.sdf, v.nx,
                     v.ny, v.nz, v.n, *v.W);
  if (hipGetLastError() != hipSuccess) return -1;
  if (hipFreeAsync(dp, v.stream) != hipSuccess) return -1;
  return project_info(v, info8);
}

}  // namespace wt

// ---------------------------------------------------------------------------------------------
// Mesh voxelizer (src/voxelize/wgslVoxel.ts:33-322, host sequence VoxelizerGPU.ts:48-151):
//   1. signed ray crossings per triangle and axis into the first cell above the crossing,
//   2. per-column prefix sums along each axis -> one vote bit per axis where the winding is non-zero,
//   3. thin surface shell (bit 8) for parts too thin to own an interior cell,
//   4. flags = 2-of-3 vote, or shell-only cells without a solid 6-neighbour,
//   5. unsigned near-wall distance (atomic min through ~bits), 6. wheel tags, 7. projection.
// This file is compiled without floating-point contraction so that the CPU restatement in the
// oracle reproduces every comparison exactly.
// ---------------------------------------------------------------------------------------------
namespace wt {
namespace {

struct MeshCtx {
  float M[16];
  int nx, ny, nz;
  int64_t n;
  int ntri;
};

__device__ inline float3 vtx(const float* __restrict__ pos, const uint32_t* __restrict__ idx, int64_t i, const MeshCtx& C) {
  const uint32_t k = idx[i];
  const float x = pos[3 * k], y = pos[3 * k + 1], z = pos[3 * k + 2];
  // column-major M: p' = M[0..3] x + M[4..7] y + M[8..11] z + M[12..15]
  return make_float3(((C.M[0] * x + C.M[4] * y) + C.M[8] * z) + C.M[12], ((C.M[1] * x + C.M[5] * y) + C.M[9] * z) + C.M[13],
                     ((C.M[2] * x + C.M[6] * y) + C.M[10] * z) + C.M[14]);
}
__device__ inline float comp(float3 p, int a) { return a == 0 ? p.x : (a == 1 ? p.y : p.z); }
__device__ inline int dim_of(const MeshCtx& C, int a) { return a == 0 ? C.nx : (a == 1 ? C.ny : C.nz); }
__device__ inline int64_t cell_of(const MeshCtx& C, int x, int y, int z) { return (int64_t)x + (int64_t)C.nx * ((int64_t)y + (int64_t)C.ny * z); }
__device__ inline float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline float dot3(float3 a, float3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ inline float len3s(float3 a) { return __builtin_sqrtf(dot3(a, a)); }
__device__ inline float3 cross3(float3 a, float3 b) {
  return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// 1. crossings (wgslVoxel.ts:33-94)
__global__ __launch_bounds__(64) void k_crossings(const float* __restrict__ pos, const uint32_t* __restrict__ idx, MeshCtx C,
                                                  int* __restrict__ hits) {
  const int t = blockIdx.x * 64 + threadIdx.x;
  if (t >= C.ntri) return;
  const float3 p0 = vtx(pos, idx, 3 * (int64_t)t, C), p1 = vtx(pos, idx, 3 * (int64_t)t + 1, C), p2 = vtx(pos, idx, 3 * (int64_t)t + 2, C);
  for (int a = 0; a < 3; a++) {
    const int ua = (a + 1) % 3, va = (a + 2) % 3;
    const float Ax = comp(p0, ua), Ay = comp(p0, va), Bx = comp(p1, ua), By = comp(p1, va), Cx = comp(p2, ua), Cy = comp(p2, va);
    const float area = (Bx - Ax) * (Cy - Ay) - (By - Ay) * (Cx - Ax);
    if (fabsf(area) < 1e-12f) continue;
    const int sgn = area > 0.0f ? -1 : 1;
    const int nu = dim_of(C, ua), nv = dim_of(C, va), na = dim_of(C, a);
    const float lox = fminf(fminf(Ax, Bx), Cx), hix = fmaxf(fmaxf(Ax, Bx), Cx);
    const float loy = fminf(fminf(Ay, By), Cy), hiy = fmaxf(fmaxf(Ay, By), Cy);
    const int u0 = max((int)ceilf(lox - 0.5f), 0), u1 = min((int)floorf(hix - 0.5f), nu - 1);
    const int v0 = max((int)ceilf(loy - 0.5f), 0), v1 = min((int)floorf(hiy - 0.5f), nv - 1);
    const float ha = comp(p0, a), hb = comp(p1, a), hc = comp(p2, a);
    const float inv = __fdiv_rn(1.0f, area);
    for (int iu = u0; iu <= u1; iu++) {
      for (int iv = v0; iv <= v1; iv++) {
        const float Px = ((float)iu + 0.5f) + 1.37e-4f, Py = ((float)iv + 0.5f) + 2.71e-4f;
        const float w0 = ((Bx - Px) * (Cy - Py) - (By - Py) * (Cx - Px)) * inv;
        const float w1 = ((Cx - Px) * (Ay - Py) - (Cy - Py) * (Ax - Px)) * inv;
        const float w2 = (1.0f - w0) - w1;
        if (w0 < 0.0f || w1 < 0.0f || w2 < 0.0f) continue;
        const float h = (w0 * ha + w1 * hb) + w2 * hc;
        int k = (int)floorf(h - 0.5f) + 1;
        if (k >= na) continue;
        k = max(k, 0);
        int64_t c;
        if (a == 0) c = cell_of(C, k, iu, iv);
        else if (a == 1) c = cell_of(C, iv, k, iu);
        else c = cell_of(C, iu, iv, k);
        atomicAdd(&hits[(int64_t)a * C.n + c], sgn);
      }
    }
  }
}

// 2. column scans (wgslVoxel.ts:96-124)
__global__ __launch_bounds__(64) void k_scan(const int* __restrict__ hits, uint32_t* __restrict__ vote, MeshCtx C, int a) {
  const int ua = (a + 1) % 3, va = (a + 2) % 3;
  const int nu = dim_of(C, ua), nv = dim_of(C, va), na = dim_of(C, a);
  const int64_t col = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (col >= (int64_t)nu * nv) return;
  const int iu = (int)(col % nu), iv = (int)(col / nu);
  int s = 0;
  for (int k = 0; k < na; k++) {
    int64_t c;
    if (a == 0) c = cell_of(C, k, iu, iv);
    else if (a == 1) c = cell_of(C, iv, k, iu);
    else c = cell_of(C, iu, iv, k);
    s += hits[(int64_t)a * C.n + c];
    if (s != 0) atomicOr(&vote[c], 1u << a);
  }
}

// 3. surface shell (wgslVoxel.ts:126-170)
__global__ __launch_bounds__(64) void k_surface(const float* __restrict__ pos, const uint32_t* __restrict__ idx, MeshCtx C,
                                                uint32_t* __restrict__ vote) {
  const int t = blockIdx.x * 64 + threadIdx.x;
  if (t >= C.ntri) return;
  const float3 p0 = vtx(pos, idx, 3 * (int64_t)t, C);
  const float3 e1 = sub3(vtx(pos, idx, 3 * (int64_t)t + 1, C), p0);
  const float3 e2 = sub3(vtx(pos, idx, 3 * (int64_t)t + 2, C), p0);
  const float m = fmaxf(fmaxf(len3s(e1), len3s(e2)), len3s(sub3(e2, e1)));
  const int steps = (int)fminf(fmaxf(ceilf(__fdiv_rn(m, 0.4f)), 1.0f), 384.0f);
  const float fs = (float)steps;
  const float3 cr = cross3(e1, e2);
  const float cl = len3s(cr);
  if (cl < 1e-12f) return;
  const float3 nrm = make_float3(__fdiv_rn(cr.x, cl), __fdiv_rn(cr.y, cl), __fdiv_rn(cr.z, cl));
  const float tol = 0.5f * fmaxf(fmaxf(fabsf(nrm.x), fabsf(nrm.y)), fabsf(nrm.z));
  const bool is2D = C.nz == 1;
  for (int i = 0; i <= steps; i++) {
    for (int j = 0; j <= steps - i; j++) {
      const float si = __fdiv_rn((float)i, fs), sj = __fdiv_rn((float)j, fs);
      const float3 p = make_float3((p0.x + e1.x * si) + e2.x * sj, (p0.y + e1.y * si) + e2.y * sj, (p0.z + e1.z * si) + e2.z * sj);
      if (p.x < 0.0f || p.y < 0.0f || p.z < 0.0f || p.x >= (float)C.nx || p.y >= (float)C.ny || p.z >= (float)C.nz) continue;
      const float cx = floorf(p.x), cy = floorf(p.y), cz = floorf(p.z);
      float3 ctr = make_float3(cx + 0.5f, cy + 0.5f, cz + 0.5f);
      float tt = tol;
      if (is2D) {
        // the 2D grid is one plane: compare the in-plane distance of the section line
        const float n2 = __builtin_sqrtf(nrm.x * nrm.x + nrm.y * nrm.y);
        if (n2 < 1e-4f) continue;
        ctr.z = p.z;
        tt = 0.5f * fmaxf(fabsf(nrm.x), fabsf(nrm.y));
        if (fabsf(p.z - 0.5f) > 0.5f) continue;
      }
      if (fabsf(dot3(sub3(ctr, p0), nrm)) > tt) continue;
      atomicOr(&vote[cell_of(C, (int)cx, (int)cy, (int)cz)], 8u);
    }
  }
}

__device__ inline bool votes(uint32_t v) { return __popc(v & 7u) >= 2; }

// 4. vote -> flags (wgslVoxel.ts:172-199)
__global__ __launch_bounds__(256) void k_finalize(const uint32_t* __restrict__ vote, uint32_t* __restrict__ flags, MeshCtx C) {
  const int64_t id = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (id >= C.n) return;
  const uint32_t v = vote[id];
  bool solid = votes(v);
  if (!solid && (v & 8u)) {
    const int x = (int)(id % C.nx), y = (int)((id / C.nx) % C.ny), z = (int)(id / ((int64_t)C.nx * C.ny));
    bool near = false;
    if (x > 0) near = near || votes(vote[id - 1]);
    if (x + 1 < C.nx) near = near || votes(vote[id + 1]);
    if (y > 0) near = near || votes(vote[id - C.nx]);
    if (y + 1 < C.ny) near = near || votes(vote[id + C.nx]);
    if (z > 0) near = near || votes(vote[id - (int64_t)C.nx * C.ny]);
    if (z + 1 < C.nz) near = near || votes(vote[id + (int64_t)C.nx * C.ny]);
    solid = !near;
  }
  flags[id] = solid ? 1u : 0u;
}

// closest point on a triangle (Ericson, Real-Time Collision Detection 5.1.5), as wgslVoxel.ts:211-227
__device__ float3 closest_pt(float3 p, float3 a, float3 b, float3 c) {
  const float3 ab = sub3(b, a), ac = sub3(c, a), ap = sub3(p, a);
  const float d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return a;
  const float3 bp = sub3(p, b);
  const float d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return b;
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    const float v = __fdiv_rn(d1, d1 - d3);
    return make_float3(a.x + ab.x * v, a.y + ab.y * v, a.z + ab.z * v);
  }
  const float3 cp = sub3(p, c);
  const float d5 = dot3(ab, cp), d6 = dot3(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return c;
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    const float w = __fdiv_rn(d2, d2 - d6);
    return make_float3(a.x + ac.x * w, a.y + ac.y * w, a.z + ac.z * w);
  }
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    const float w = __fdiv_rn(d4 - d3, (d4 - d3) + (d5 - d6));
    return make_float3(b.x + (c.x - b.x) * w, b.y + (c.y - b.y) * w, b.z + (c.z - b.z) * w);
  }
  const float den = __fdiv_rn(1.0f, (va + vb) + vc);
  const float v = vb * den, w = vc * den;
  return make_float3((a.x + ab.x * v) + ac.x * w, (a.y + ab.y * v) + ac.y * w, (a.z + ab.z * v) + ac.z * w);
}

// 5. near-wall distances (wgslVoxel.ts:205-265): atomic max of ~bits = atomic min of the distance
__global__ __launch_bounds__(64) void k_mesh_sdf(const float* __restrict__ pos, const uint32_t* __restrict__ idx, MeshCtx C,
                                                 uint32_t* __restrict__ sdf) {
  const int t = blockIdx.x * 64 + threadIdx.x;
  if (t >= C.ntri) return;
  const float3 a = vtx(pos, idx, 3 * (int64_t)t, C), b = vtx(pos, idx, 3 * (int64_t)t + 1, C), c = vtx(pos, idx, 3 * (int64_t)t + 2, C);
  const float3 e1 = sub3(b, a), e2 = sub3(c, a);
  if (len3s(cross3(e1, e2)) < 1e-12f) return;
  const float m = fmaxf(fmaxf(len3s(e1), len3s(e2)), len3s(sub3(e2, e1)));
  const int steps = (int)fminf(fmaxf(ceilf(__fdiv_rn(m, 0.7f)), 1.0f), 256.0f);
  const float fs = (float)steps;
  const bool is2D = C.nz == 1;
  const int R = 2;
  for (int i = 0; i <= steps; i++) {
    for (int j = 0; j <= steps - i; j++) {
      const float si = __fdiv_rn((float)i, fs), sj = __fdiv_rn((float)j, fs);
      const float3 p = make_float3((a.x + e1.x * si) + e2.x * sj, (a.y + e1.y * si) + e2.y * sj, (a.z + e1.z * si) + e2.z * sj);
      if (is2D && fabsf(p.z - 0.5f) > 2.5f) continue;
      const int bx = (int)floorf(p.x), by = (int)floorf(p.y), bz = (int)floorf(p.z);
      for (int dz = -R; dz <= R; dz++) {
        if (is2D && dz != 0) continue;
        for (int dy = -R; dy <= R; dy++) {
          for (int dx = -R; dx <= R; dx++) {
            const int qx = bx + dx, qy = by + dy, qz = is2D ? 0 : bz + dz;
            if (qx < 0 || qy < 0 || qz < 0 || qx >= C.nx || qy >= C.ny || qz >= C.nz) continue;
            const float3 ctr = make_float3((float)qx + 0.5f, (float)qy + 0.5f, (float)qz + 0.5f);
            const float d = len3s(sub3(ctr, closest_pt(ctr, a, b, c)));
            if (d > 2.5f) continue;
            atomicMax(&sdf[cell_of(C, qx, qy, qz)], ~__float_as_uint(d));
          }
        }
      }
    }
  }
}

// 6. wheel tags (wgslVoxel.ts:268-290)
__global__ __launch_bounds__(256) void k_wheel_tag(uint32_t* __restrict__ flags, MeshCtx C, KWheels W) {
  const int64_t id = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (id >= C.n || flags[id] == 0u) return;
  const float3 p = make_float3((float)(id % C.nx) + 0.5f, (float)((id / C.nx) % C.ny) + 0.5f, (float)(id / ((int64_t)C.nx * C.ny)) + 0.5f);
  for (int k = 0; k < W.count; k++) {
    const float4 a = W.w[2 * k], b = W.w[2 * k + 1];
    const float3 d = make_float3(p.x - a.x, p.y - a.y, p.z - a.z);
    const float al = dot3(d, make_float3(b.x, b.y, b.z));
    const float rad = len3s(make_float3(d.x - al * b.x, d.y - al * b.y, d.z - al * b.z));
    if (fabsf(al) <= b.w + 0.6f && rad <= __fdiv_rn(1.0f, a.w) + 0.75f) {
      flags[id] = 1u | ((uint32_t)(k + 1) << 8);
      return;
    }
  }
}

}  // namespace

int project_info(const VoxCtx& v, int64_t* info8);

int voxelize_mesh(const VoxCtx& v, const float* positions, int nv, const uint32_t* indices, int nt, const float* M16,
                  int64_t* info8) {
  hipStream_t st = v.stream;
  // an empty mesh clears flags and distances (VoxelizerGPU.ts:54-58)
  if (nt == 0) {
    if (hipMemsetAsync(v.flags, 0, 4 * v.n, st) != hipSuccess) return -1;
    if (hipMemsetAsync(v.sdf, 0, 4 * v.n, st) != hipSuccess) return -1;
    for (int k = 0; k < 8; k++) info8[k] = 0;
    return hipStreamSynchronize(st) == hipSuccess ? 0 : -1;
  }
  MeshCtx C;
  for (int k = 0; k < 16; k++) C.M[k] = M16[k];
  C.nx = v.nx; C.ny = v.ny; C.nz = v.nz; C.n = v.n; C.ntri = nt;
  float* dpos = nullptr;
  uint32_t* didx = nullptr;
  int* hits = nullptr;
  uint32_t* vote = nullptr;
  auto fail = [&]() {
    if (dpos) (void)hipFreeAsync(dpos, st);
    if (didx) (void)hipFreeAsync(didx, st);
    if (hits) (void)hipFreeAsync(hits, st);
    if (vote) (void)hipFreeAsync(vote, st);
    (void)hipStreamSynchronize(st);
    return -1;
  };
  if (hipMallocAsync((void**)&dpos, sizeof(float) * 3 * (size_t)nv, st) != hipSuccess) return fail();
  if (hipMallocAsync((void**)&didx, sizeof(uint32_t) * 3 * (size_t)nt, st) != hipSuccess) return fail();
  if (hipMallocAsync((void**)&hits, sizeof(int) * 3 * (size_t)v.n, st) != hipSuccess) return fail();
  if (hipMallocAsync((void**)&vote, sizeof(uint32_t) * (size_t)v.n, st) != hipSuccess) return fail();
  if (hipMemcpyAsync(dpos, positions, sizeof(float) * 3 * (size_t)nv, hipMemcpyHostToDevice, st) != hipSuccess) return fail();
  if (hipMemcpyAsync(didx, indices, sizeof(uint32_t) * 3 * (size_t)nt, hipMemcpyHostToDevice, st) != hipSuccess) return fail();
  if (hipMemsetAsync(hits, 0, sizeof(int) * 3 * (size_t)v.n, st) != hipSuccess) return fail();
  if (hipMemsetAsync(vote, 0, sizeof(uint32_t) * (size_t)v.n, st) != hipSuccess) return fail();
  if (hipMemsetAsync(v.sdf, 0, sizeof(uint32_t) * (size_t)v.n, st) != hipSuccess) return fail();
  const unsigned tb = (unsigned)((nt + 63) / 64);
  hipLaunchKernelGGL(k_crossings, dim3(tb), dim3(64), 0, st, dpos, didx, C, hits);
  const int dims[3] = {v.nx, v.ny, v.nz};
  for (int a = 0; a < 3; a++) {
    const int64_t cols = (int64_t)dims[(a + 1) % 3] * dims[(a + 2) % 3];
    hipLaunchKernelGGL(k_scan, dim3((unsigned)((cols + 63) / 64)), dim3(64), 0, st, hits, vote, C, a);
  }
  hipLaunchKernelGGL(k_surface, dim3(tb), dim3(64), 0, st, dpos, didx, C, vote);
  hipLaunchKernelGGL(k_finalize, dim3((unsigned)((v.n + 255) / 256)), dim3(256), 0, st, vote, v.flags, C);
  hipLaunchKernelGGL(k_mesh_sdf, dim3(tb), dim3(64), 0, st, dpos, didx, C, v.sdf);
  if (v.W->count > 0) hipLaunchKernelGGL(k_wheel_tag, dim3((unsigned)((v.n + 255) / 256)), dim3(256), 0, st, v.flags, C, *v.W);
  if (hipGetLastError() != hipSuccess) return fail();
  (void)hipFreeAsync(dpos, st);
  (void)hipFreeAsync(didx, st);
  (void)hipFreeAsync(hits, st);
  (void)hipFreeAsync(vote, st);
  return project_info(v, info8);
}

}  // namespace wt
