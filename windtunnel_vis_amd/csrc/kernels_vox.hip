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

namespace wt {
int voxelize_mesh(const VoxCtx&, const float*, int, const uint32_t*, int, const float*, int64_t*) { return -1; }
}  // namespace wt
