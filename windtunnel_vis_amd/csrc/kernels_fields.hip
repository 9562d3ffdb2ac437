// Post-processing passes outside the step path: derived fields (vorticity, Q-criterion,
// turbulence intensity, wgslLbm.ts:527-592) and the strided field export (SolverGPU.ts:374-425).
// All fields are one 8-byte record of four IEEE halves per cell (the reference's rgba16float
// textures), x fastest, so the export reads exactly the values the reference's textures hold.
#include "launch.hpp"

namespace wt {
namespace {

constexpr int FBLOCK = 256;

__device__ inline float4 ld_h4(const uint2* __restrict__ p, int64_t i) {
  const uint2 u = p[i];
  return make_float4((float)__builtin_bit_cast(_Float16, (uint16_t)(u.x & 0xffffu)),
                     (float)__builtin_bit_cast(_Float16, (uint16_t)(u.x >> 16)),
                     (float)__builtin_bit_cast(_Float16, (uint16_t)(u.y & 0xffffu)),
                     (float)__builtin_bit_cast(_Float16, (uint16_t)(u.y >> 16)));
}

__device__ inline uint2 pack_h4(float a, float b, float c, float d) {  // round to nearest even
  const uint32_t ha = __builtin_bit_cast(uint16_t, (_Float16)a), hb = __builtin_bit_cast(uint16_t, (_Float16)b);
  const uint32_t hc = __builtin_bit_cast(uint16_t, (_Float16)c), hd = __builtin_bit_cast(uint16_t, (_Float16)d);
  return make_uint2(ha | (hb << 16), hc | (hd << 16));
}

// One thread per cell. The velocity record of a cell is (ux, uy, uz, rho) from the macro pass;
// neighbour reads clamp at the domain faces (textureLoad of a clamped coordinate), solid tests
// read false outside the domain (isSolidAt).
template <int DIM>
__global__ __launch_bounds__(FBLOCK) void k_derived(const uint2* __restrict__ vel, const uint8_t* __restrict__ kind,
                                                    const float4* __restrict__ meanA, const float* __restrict__ meanB,
                                                    uint2* __restrict__ vort, uint2* __restrict__ stat, uint2* __restrict__ mean,
                                                    KParams P) {
  const int64_t idx = (int64_t)blockIdx.x * FBLOCK + threadIdx.x;
  if (idx >= P.n) return;
  const uint32_t row = P.divx.div((uint32_t)idx);
  const int ix = (int)((uint32_t)idx - row * (uint32_t)P.nx);
  const uint32_t zz = P.divy.div(row);
  const int iy = (int)(row - zz * (uint32_t)P.ny), iz = (int)zz;
  const int NX = P.nx, NY = P.ny, NZ = P.nz;
  const int64_t sx = 1, sy = NX, sz = (int64_t)NX * NY;
  const bool solid = (kind[idx] & KIND_SOLID) != 0;
  float4 vo = make_float4(0.f, 0.f, 0.f, 0.f);
  if (!solid) {
    // central differences, one-sided (halved) at a clamped face
    const int64_t xp = ix + 1 < NX ? idx + sx : idx, xm = ix > 0 ? idx - sx : idx;
    const int64_t yp = iy + 1 < NY ? idx + sy : idx, ym = iy > 0 ? idx - sy : idx;
    const float4 ax = ld_h4(vel, xp), bx = ld_h4(vel, xm), ay = ld_h4(vel, yp), by = ld_h4(vel, ym);
    // g[j][i] = d u_i / d x_j
    float g[3][3] = {{0.5f * (ax.x - bx.x), 0.5f * (ax.y - bx.y), 0.5f * (ax.z - bx.z)},
                     {0.5f * (ay.x - by.x), 0.5f * (ay.y - by.y), 0.5f * (ay.z - by.z)},
                     {0.f, 0.f, 0.f}};
    if (DIM == 3) {
      const int64_t zp = iz + 1 < NZ ? idx + sz : idx, zm = iz > 0 ? idx - sz : idx;
      const float4 az = ld_h4(vel, zp), bz = ld_h4(vel, zm);
      g[2][0] = 0.5f * (az.x - bz.x);
      g[2][1] = 0.5f * (az.y - bz.y);
      g[2][2] = 0.5f * (az.z - bz.z);
    }
    const float wx = g[1][2] - g[2][1], wy = g[2][0] - g[0][2], wz = g[0][1] - g[1][0];
    float S2 = 0.f, W2 = 0.f;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const float s = 0.5f * (g[j][i] + g[i][j]), w = 0.5f * (g[j][i] - g[i][j]);
        S2 += s * s;
        W2 += w * w;
      }
    // Q fades out within two cells of a solid along the axes (1 cell: off, 2 cells: half)
    auto solid_at = [&](int x, int y, int z) -> bool {
      if (x < 0 || y < 0 || z < 0 || x >= NX || y >= NY || z >= NZ) return false;
      return (kind[(int64_t)x + sy * y + sz * z] & KIND_SOLID) != 0;
    };
    float near = 0.f;
#pragma unroll
    for (int d = 1; d <= 2; d++) {
      const float v = d == 1 ? 1.0f : 0.5f;
      bool hit = solid_at(ix + d, iy, iz) || solid_at(ix - d, iy, iz) || solid_at(ix, iy + d, iz) || solid_at(ix, iy - d, iz);
      if (DIM == 3) hit = hit || solid_at(ix, iy, iz + d) || solid_at(ix, iy, iz - d);
      if (hit) near = fmaxf(near, v);
    }
    vo = make_float4(wx, wy, wz, 0.5f * (W2 - S2) * (1.0f - near));
  }
  const float4 m = meanA[idx];
  const float k = fmaxf(meanB[idx] - (m.x * m.x + m.y * m.y + m.z * m.z), 0.0f);
  const float ti = sqrtf(k / 3.0f) / fmaxf(P.uRef, 1e-6f);
  vort[idx] = pack_h4(vo.x, vo.y, vo.z, vo.w);
  stat[idx] = pack_h4(solid ? 0.0f : ti, solid ? 1.0f : 0.0f, 0.0f, 0.0f);
  mean[idx] = pack_h4(m.x, m.y, m.z, m.w);
}

// Output point g samples cell min(g * stride, n - 1) per axis; three float4 arrays, x fastest.
__global__ __launch_bounds__(FBLOCK) void k_export(const uint2* __restrict__ vel, const uint2* __restrict__ mean,
                                                   const uint2* __restrict__ vort, int nx, int ny, int nz, int stride, int ox,
                                                   int oy, int64_t nout, float4* __restrict__ ov, float4* __restrict__ om,
                                                   float4* __restrict__ ow) {
  const int64_t i = (int64_t)blockIdx.x * FBLOCK + threadIdx.x;
  if (i >= nout) return;
  const int64_t gx = i % ox, r = i / ox, gy = r % oy, gz = r / oy;
  const int64_t px = min(gx * stride, (int64_t)nx - 1), py = min(gy * stride, (int64_t)ny - 1),
                pz = min(gz * stride, (int64_t)nz - 1);
  const int64_t c = px + (int64_t)nx * (py + (int64_t)ny * pz);
  ov[i] = ld_h4(vel, c);
  om[i] = ld_h4(mean, c);
  ow[i] = ld_h4(vort, c);
}

// Half records to float4 (full-resolution readback).
__global__ __launch_bounds__(FBLOCK) void k_unpack_h4(const uint2* __restrict__ in, int64_t n, float4* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * FBLOCK + threadIdx.x;
  if (i < n) out[i] = ld_h4(in, i);
}

inline unsigned fblocks(int64_t n) { return (unsigned)((n + FBLOCK - 1) / FBLOCK); }

}  // namespace

void launch_derived(const LaunchCtx& c, const void* vel, const void* meanA, const float* meanB, void* vort, void* stat, void* mean) {
  if (c.P.nz > 1)
    hipLaunchKernelGGL((k_derived<3>), dim3(fblocks(c.P.n)), dim3(FBLOCK), 0, c.stream, (const uint2*)vel, c.kind,
                       (const float4*)meanA, meanB, (uint2*)vort, (uint2*)stat, (uint2*)mean, c.P);
  else
    hipLaunchKernelGGL((k_derived<2>), dim3(fblocks(c.P.n)), dim3(FBLOCK), 0, c.stream, (const uint2*)vel, c.kind,
                       (const float4*)meanA, meanB, (uint2*)vort, (uint2*)stat, (uint2*)mean, c.P);
}

void launch_export(hipStream_t st, const void* vel, const void* mean, const void* vort, int nx, int ny, int nz, int stride,
                   const int (&o)[3], float* ov, float* om, float* ow) {
  const int64_t nout = (int64_t)o[0] * o[1] * o[2];
  hipLaunchKernelGGL(k_export, dim3(fblocks(nout)), dim3(FBLOCK), 0, st, (const uint2*)vel, (const uint2*)mean, (const uint2*)vort,
                     nx, ny, nz, stride, o[0], o[1], nout, (float4*)ov, (float4*)om, (float4*)ow);
}

void launch_unpack_h4(hipStream_t st, const void* in, int64_t n, float* out) {
  hipLaunchKernelGGL(k_unpack_h4, dim3(fblocks(n)), dim3(FBLOCK), 0, st, (const uint2*)in, n, (float4*)out);
}

}  // namespace wt
