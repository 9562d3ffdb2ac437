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
  const uint8_t kd = kind[idx];
  const bool solid = (kd & KIND_SOLID) != 0;
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
    // Q fades out within two cells of a solid along the axes (1 cell: off, 2 cells: half); the
    // classification pass marks those cells (KIND_NEAR1 / KIND_NEAR2) once per geometry
    const float near = (kd & KIND_NEAR1) ? 1.0f : (kd & KIND_NEAR2) ? 0.5f : 0.0f;
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

// ---- field sampling, probes, streamlines and particle tracers (render/tracers.ts) -------------
// A field record is sampled like the reference's linear-filtered, clamp-to-edge texture: texel
// centres at cell centres (i + 0.5), trilinear in FP32 (the reference's filter precision is the
// texture unit's); in 2D (nz == 1) the single layer.
struct FieldView {
  const uint2* f;
  int nx, ny, nz;
};

__device__ inline float4 sample_h4(const FieldView& v, float px, float py, float pz) {
  const bool d2 = v.nz == 1;
  const float tx = px - 0.5f, ty = py - 0.5f, tz = d2 ? 0.0f : pz - 0.5f;
  const float fx0 = floorf(tx), fy0 = floorf(ty), fz0 = floorf(tz);
  const float ax = tx - fx0, ay = ty - fy0, az = tz - fz0;
  auto cl = [](float c, int n) { return (int64_t)(int)fminf(fmaxf(c, 0.0f), (float)(n - 1)); };  // NaN -> 0
  const int64_t x0 = cl(fx0, v.nx), x1 = cl(fx0 + 1.0f, v.nx), y0 = cl(fy0, v.ny), y1 = cl(fy0 + 1.0f, v.ny);
  const int64_t z0 = cl(fz0, v.nz), z1 = cl(fz0 + 1.0f, v.nz);
  const int64_t sy = v.nx, sz = (int64_t)v.nx * v.ny;
  auto lerp4 = [](float4 a, float4 b, float t) {
    return make_float4(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t, a.w + (b.w - a.w) * t);
  };
  const float4 c00 = lerp4(ld_h4(v.f, x0 + y0 * sy + z0 * sz), ld_h4(v.f, x1 + y0 * sy + z0 * sz), ax);
  const float4 c10 = lerp4(ld_h4(v.f, x0 + y1 * sy + z0 * sz), ld_h4(v.f, x1 + y1 * sy + z0 * sz), ax);
  const float4 c01 = lerp4(ld_h4(v.f, x0 + y0 * sy + z1 * sz), ld_h4(v.f, x1 + y0 * sy + z1 * sz), ax);
  const float4 c11 = lerp4(ld_h4(v.f, x0 + y1 * sy + z1 * sz), ld_h4(v.f, x1 + y1 * sy + z1 * sz), ax);
  return lerp4(lerp4(c00, c10, ay), lerp4(c01, c11, ay), az);
}

__device__ inline float3 sample_vel(const FieldView& v, float3 p) {
  const float4 s = sample_h4(v, p.x, p.y, p.z);
  return make_float3(s.x, s.y, v.nz == 1 ? 0.0f : s.z);
}

// solid test of the cell holding p (false outside the domain)
__device__ inline bool solid_at(const uint8_t* __restrict__ kind, int nx, int ny, int nz, float3 p) {
  const int cx = (int)floorf(p.x), cy = (int)floorf(p.y), cz = nz == 1 ? 0 : (int)floorf(p.z);
  if (cx < 0 || cy < 0 || cz < 0 || cx >= nx || cy >= ny || cz >= nz) return false;
  return (kind[(int64_t)cx + (int64_t)nx * ((int64_t)cy + (int64_t)ny * cz)] & KIND_SOLID) != 0;
}

__global__ __launch_bounds__(64) void k_probes(FieldView vel, FieldView mean, const float4* __restrict__ pos, int count,
                                               float4* __restrict__ out) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= count) return;
  const float4 p = pos[i];
  float4 v = sample_h4(vel, p.x, p.y, p.z), m = sample_h4(mean, p.x, p.y, p.z);
  if (vel.nz == 1) { v.z = 0.0f; m.z = 0.0f; }
  out[2 * i] = v;
  out[2 * i + 1] = m;
}

__device__ inline float3 f3(const float (&a)[4]) { return make_float3(a[0], a[1], a[2]); }

// Streamlines from a rake: seed s sits on line s / perLine, at (s % perLine + 0.5) / perLine along
// a -> b, offset by (line - (lines - 1) / 2) d; each step goes h along the normalised velocity at
// the midpoint; a line stops (w = -1 from then on) outside the domain, in a solid or below 1e-3 U.
__global__ __launch_bounds__(64) void k_streamlines(FieldView vel, const uint8_t* __restrict__ kind, wt_rake rake, int seeds,
                                                    int per_line, int points, float h, float U, float4* __restrict__ out) {
  const int s = blockIdx.x * 64 + threadIdx.x;
  if (s >= seeds) return;
  const int per = max(per_line, 1);
  const float l = (float)(s / per) - 0.5f * (float)(seeds / per - 1);
  const float t = ((float)(s % per) + 0.5f) / (float)per;
  const float3 a = f3(rake.a), b = f3(rake.b), d = f3(rake.d);
  float3 p = make_float3(a.x + (b.x - a.x) * t + d.x * l, a.y + (b.y - a.y) * t + d.y * l, a.z + (b.z - a.z) * t + d.z * l);
  const bool d2 = vel.nz == 1;
  if (d2) p.z = 0.5f;
  const int nx = vel.nx, ny = vel.ny, nz = vel.nz;
  bool alive = true;
  const float Umin = fmaxf(U, 1e-5f) * 1e-3f;
  for (int k = 0; k < points; k++) {
    const float3 v = sample_vel(vel, p);
    const float sp = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    const bool out_of = d2 ? (p.x < 0.f || p.y < 0.f || p.x >= nx || p.y >= ny)
                           : (p.x < 0.f || p.y < 0.f || p.z < 0.f || p.x >= nx || p.y >= ny || p.z >= nz);
    if (alive && (out_of || solid_at(kind, nx, ny, nz, p) || sp < Umin)) alive = false;
    out[(int64_t)s * points + k] = make_float4(p.x, p.y, p.z, alive ? sp : -1.0f);
    if (!alive) continue;
    const float3 dir = make_float3(v.x / sp, v.y / sp, v.z / sp);
    const float3 vm = sample_vel(vel, make_float3(p.x + dir.x * 0.5f * h, p.y + dir.y * 0.5f * h, p.z + dir.z * 0.5f * h));
    const float lm = sqrtf(vm.x * vm.x + vm.y * vm.y + vm.z * vm.z);
    if (lm < 1e-9f) { alive = false; continue; }
    p = make_float3(p.x + vm.x / lm * h, p.y + vm.y / lm * h, p.z + vm.z / lm * h);
  }
}

// PCG-RXS-M-XS 32-bit hash (Jarzynski & Olano, "Hash Functions for GPU Rendering", JCGT 2020)
__device__ inline uint32_t pcg_hash(uint32_t v) {
  const uint32_t s = v * 747796405u + 2891336453u;
  const uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
  return (w >> 22u) ^ w;
}

struct Rng {
  uint32_t s;
  __device__ float next() {
    s = pcg_hash(s);
    return (float)s / 4294967295.0f;
  }
};

struct ParticleArgs {
  int count, trail, head, emitter;  // emitter 0: inlet plane, 1: rake nozzles
  float steps, max_age;
  uint32_t seed;
  int nozzles, lines, fill, advance, pad;
  wt_rake rake;
};

__device__ inline float3 spawn(Rng& r, const ParticleArgs& A, int nx, int ny, int nz) {
  const bool d2 = nz == 1;
  if (A.emitter == 0) {
    const float x = 1.0f + r.next() * 2.0f, y = 0.5f + r.next() * (float)(ny - 1);
    const float z = d2 ? 0.5f : 0.5f + r.next() * (float)(nz - 1);
    return make_float3(x, y, z);
  }
  const uint32_t lines = (uint32_t)max(A.lines, 1), noz = (uint32_t)max(A.nozzles, 1);
  const uint32_t k = min((uint32_t)(r.next() * (float)(noz * lines)), noz * lines - 1u);
  const float t = ((float)(k % noz) + 0.5f) / (float)noz;
  const float l = (float)(k / noz) - 0.5f * (float)(lines - 1u);
  const float3 a = f3(A.rake.a), b = f3(A.rake.b), d = f3(A.rake.d);
  float3 p = make_float3(a.x + (b.x - a.x) * t + d.x * l, a.y + (b.y - a.y) * t + d.y * l, a.z + (b.z - a.z) * t + d.z * l);
  const float jx = r.next() - 0.5f, jy = r.next() - 0.5f, jz = r.next() - 0.5f;
  p = make_float3(p.x + jx * 0.35f, p.y + jy * 0.35f, p.z + jz * 0.35f);
  if (d2) p.z = 0.5f;
  return p;
}

// One frame of the particle tracers: fill (fresh positions, the whole trail set to them), or age
// waiting particles, or advect with RK2 (midpoint) in sub-steps of at most 0.6 cells, respawning
// particles that leave, hit a solid or outlive 4 x maxAge; the new position goes to trail slot head.
__global__ __launch_bounds__(128) void k_particles(FieldView vel, const uint8_t* __restrict__ kind, ParticleArgs A,
                                                   float4* __restrict__ state, float4* __restrict__ hist) {
  const int i = blockIdx.x * 128 + threadIdx.x;
  if (i >= A.count) return;
  const int nx = vel.nx, ny = vel.ny, nz = vel.nz;
  const bool d2 = nz == 1;
  Rng r{pcg_hash((uint32_t)i ^ pcg_hash(A.seed))};
  float4 s = state[i];
  if (A.fill) {
    if (A.emitter == 0) {
      float3 p;
      bool found = false;
      for (int k = 0; k < 8 && !found; k++) {
        const float x = 1.0f + r.next() * (float)(nx - 2), y = 0.5f + r.next() * (float)(ny - 1);
        const float z = d2 ? 0.5f : 0.5f + r.next() * (float)(nz - 1);
        p = make_float3(x, y, z);
        found = !solid_at(kind, nx, ny, nz, p);
      }
      if (!found) p = spawn(r, A, nx, ny, nz);
      s = make_float4(p.x, p.y, p.z, (float)A.trail);
    } else {
      const float3 p = spawn(r, A, nx, ny, nz);
      s = make_float4(p.x, p.y, p.z, -floorf(r.next() * A.max_age));
    }
    state[i] = s;
    for (int k = 0; k < A.trail; k++) hist[(int64_t)k * A.count + i] = make_float4(s.x, s.y, s.z, 0.0f);
    return;
  }
  if (!A.advance) return;
  if (s.w < 0.0f) {
    s.w += 1.0f;
    if (s.w >= 0.0f) {
      const float3 p = spawn(r, A, nx, ny, nz);
      s = make_float4(p.x, p.y, p.z, 0.0f);
    }
  } else {
    float3 p = make_float3(s.x, s.y, s.z);
    const float3 v0 = sample_vel(vel, p);
    const float disp = sqrtf(v0.x * v0.x + v0.y * v0.y + v0.z * v0.z) * A.steps;
    const int nsub = (int)fminf(fmaxf(ceilf(disp / 0.6f), 1.0f), 24.0f);
    const float h = A.steps / (float)nsub;
    for (int k = 0; k < nsub; k++) {
      const float3 a = sample_vel(vel, p);
      const float3 b = sample_vel(vel, make_float3(p.x + a.x * 0.5f * h, p.y + a.y * 0.5f * h, p.z + a.z * 0.5f * h));
      p = make_float3(p.x + b.x * h, p.y + b.y * h, p.z + b.z * h);
    }
    s = make_float4(p.x, p.y, p.z, s.w + 1.0f);
    const bool out_of = d2 ? (p.x < 0.f || p.y < 0.f || p.x >= nx - 1.0f || p.y >= ny)
                           : (p.x < 0.f || p.y < 0.f || p.z < 0.f || p.x >= nx - 1.0f || p.y >= ny || p.z >= nz);
    if (out_of || solid_at(kind, nx, ny, nz, p) || s.w > A.max_age * 4.0f) {
      const float3 q = spawn(r, A, nx, ny, nz);
      s = make_float4(q.x, q.y, q.z, 0.0f);
    }
  }
  state[i] = s;
  hist[(int64_t)A.head * A.count + i] = make_float4(s.x, s.y, s.z, 0.0f);
}

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

void launch_probes(hipStream_t st, const void* vel, const void* mean, const uint8_t* kind, int nx, int ny, int nz, const float* pos,
                   int count, float* out) {
  (void)kind;
  const FieldView v{(const uint2*)vel, nx, ny, nz}, m{(const uint2*)mean, nx, ny, nz};
  hipLaunchKernelGGL(k_probes, dim3((count + 63) / 64), dim3(64), 0, st, v, m, (const float4*)pos, count, (float4*)out);
}

void launch_streamlines(hipStream_t st, const void* field, const uint8_t* kind, int nx, int ny, int nz, const wt_rake& rake, int seeds,
                        int per_line, int points, float h, float U, float* out) {
  const FieldView v{(const uint2*)field, nx, ny, nz};
  hipLaunchKernelGGL(k_streamlines, dim3((seeds + 63) / 64), dim3(64), 0, st, v, kind, rake, seeds, per_line, points, h, U,
                     (float4*)out);
}

void launch_particles(hipStream_t st, const void* vel, const uint8_t* kind, int nx, int ny, int nz, const wt_particle_opts& o, int head,
                      float steps, uint32_t seed, int fill, int advance, float* state, float* hist) {
  const FieldView v{(const uint2*)vel, nx, ny, nz};
  ParticleArgs A{};
  A.count = o.count; A.trail = o.trail; A.head = head; A.emitter = o.emitter;
  A.steps = steps; A.max_age = o.max_age; A.seed = seed;
  A.nozzles = o.nozzles; A.lines = o.lines; A.fill = fill; A.advance = advance;
  A.rake = o.rake;
  hipLaunchKernelGGL(k_particles, dim3((o.count + 127) / 128), dim3(128), 0, st, v, kind, A, (float4*)state, (float4*)hist);
}

void launch_unpack_h4(hipStream_t st, const void* in, int64_t n, float* out) {
  hipLaunchKernelGGL(k_unpack_h4, dim3(fblocks(n)), dim3(FBLOCK), 0, st, (const uint2*)in, n, (float4*)out);
}

}  // namespace wt
