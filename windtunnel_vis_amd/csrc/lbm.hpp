// Device-side building blocks of the lattice-Boltzmann hot path for gfx950 (MI355X).
//
// Layout in HBM (all buffers owned by the solver handle):
//   populations  SoA, direction-major: f[i * n + cell], cell = x + nx * (y + ny * z), x fastest.
//                FP32 stores full f_i; FP16 / BF16 store the deviation f_i - w_i (RNE), so two
//                x-neighbouring cells form one 32-bit word exactly like the reference's
//                pack2x16float pairs (src/solver/wgslLbm.ts:62-75).
//   flags        u32 per cell: 0 fluid, 1 solid, 1 | (k+1) << 8 solid of wheel k (types as in
//                src/solver/SolverGPU.ts:341-344).
//   sdf          u32 per cell: ~bits(distance) or 0 = unknown (SolverGPU.ts:312-322).
//   kind         u8 per cell, derived from flags: 0 interior fluid (bulk kernel), 1 solid,
//                2 fluid that needs the boundary path, | 4 when a tyre contact patch holds the
//                belt still at this floor cell (wgslLbm.ts:276-287).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace wt {

enum : uint8_t { KIND_FLUID = 0, KIND_SOLID = 1, KIND_BOUNDARY = 2, KIND_BELT_HOLD = 4 };

// ---------------------------------------------------------------------------------------------
// Lattices (src/solver/lattice.ts:39-60). opp / mirror tables are found by search at compile
// time, as lattice.ts:20-37 finds them at run time.
// ---------------------------------------------------------------------------------------------
struct D2Q9 {
  static constexpr int Q = 9, DIM = 2;
  __host__ __device__ static constexpr int c(int i, int a) {
    constexpr int t[9][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {-1, 0, 0}, {0, -1, 0},
                             {1, 1, 0}, {-1, 1, 0}, {-1, -1, 0}, {1, -1, 0}};
    return t[i][a];
  }
  __host__ __device__ static constexpr float w(int i) { return i == 0 ? 4.0f / 9.0f : (i < 5 ? 1.0f / 9.0f : 1.0f / 36.0f); }
  __host__ __device__ static constexpr double wd(int i) { return i == 0 ? 4.0 / 9.0 : (i < 5 ? 1.0 / 9.0 : 1.0 / 36.0); }
};

struct D3Q19 {
  static constexpr int Q = 19, DIM = 3;
  __host__ __device__ static constexpr int c(int i, int a) {
    constexpr int t[19][3] = {{0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},
                              {0, 0, 1},  {0, 0, -1},  {1, 1, 0},  {-1, -1, 0}, {1, 0, 1},
                              {-1, 0, -1}, {0, 1, 1},  {0, -1, -1}, {1, -1, 0}, {-1, 1, 0},
                              {1, 0, -1}, {-1, 0, 1},  {0, 1, -1}, {0, -1, 1}};
    return t[i][a];
  }
  __host__ __device__ static constexpr float w(int i) { return i == 0 ? 1.0f / 3.0f : (i < 7 ? 1.0f / 18.0f : 1.0f / 36.0f); }
  __host__ __device__ static constexpr double wd(int i) { return i == 0 ? 1.0 / 3.0 : (i < 7 ? 1.0 / 18.0 : 1.0 / 36.0); }
};

template <class L>
__host__ __device__ constexpr int find_dir(int x, int y, int z) {
  for (int k = 0; k < L::Q; k++)
    if (L::c(k, 0) == x && L::c(k, 1) == y && L::c(k, 2) == z) return k;
  return -1;
}
template <class L> __host__ __device__ constexpr int opp(int i) { return find_dir<L>(-L::c(i, 0), -L::c(i, 1), -L::c(i, 2)); }
template <class L> __host__ __device__ constexpr int mirY(int i) { return find_dir<L>(L::c(i, 0), -L::c(i, 1), L::c(i, 2)); }
template <class L> __host__ __device__ constexpr int mirZ(int i) { return find_dir<L>(L::c(i, 0), L::c(i, 1), -L::c(i, 2)); }
template <class L> __host__ __device__ constexpr int mirYZ(int i) { return find_dir<L>(L::c(i, 0), -L::c(i, 1), -L::c(i, 2)); }

// ---------------------------------------------------------------------------------------------
// Population storage formats.
// ---------------------------------------------------------------------------------------------
typedef uint32_t u32_a2 __attribute__((aligned(2)));
typedef uint64_t u64_a4 __attribute__((aligned(4)));
typedef float f32x2 __attribute__((ext_vector_type(2), aligned(4)));

__device__ inline float h2f(uint32_t h) { return __half2float(__ushort_as_half((unsigned short)h)); }
__device__ inline uint32_t f2h(float f) { return (uint32_t)__half_as_ushort(__float2half_rn(f)); }
__device__ inline float b2f(uint32_t b) { return __uint_as_float(b << 16); }
// float -> bfloat16 bits, round to nearest even (NaN kept quiet).
__device__ inline uint32_t f2b(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40u;
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

struct SF32 {
  using T = float;
  static constexpr int BYTES = 4;
  __device__ static float ld(const T* p, int64_t e, float) { return p[e]; }
  __device__ static T enc(float f, float) { return f; }
  __device__ static T rest(float w) { return w; }
  // two consecutive cells: e is the element index of the first
  template <bool ALIGNED>
  __device__ static void ld2(const T* p, int64_t e, float, float& a, float& b) {
    // FP32 pairs are only 4-byte aligned in general (odd nx or odd n allowed)
    const uint64_t v = *(const u64_a4*)(p + e);
    a = __uint_as_float((uint32_t)v); b = __uint_as_float((uint32_t)(v >> 32));
  }
  __device__ static void st2(T* p, int64_t e, float w, float a, float b) {
    (void)w;
    f32x2 v = {a, b};
    __builtin_nontemporal_store(v, (f32x2*)(p + e));
  }
};

struct SF16 {
  using T = uint16_t;
  static constexpr int BYTES = 2;
  __device__ static float ld(const T* p, int64_t e, float w) { return h2f(p[e]) + w; }
  __device__ static T enc(float f, float w) { return (T)f2h(f - w); }
  __device__ static T enc_dev(float d) { return (T)f2h(d); }
  __device__ static T rest(float) { return 0; }
  template <bool ALIGNED>
  __device__ static void ld2(const T* p, int64_t e, float w, float& a, float& b) {
    uint32_t v = ALIGNED ? *(const uint32_t*)(p + e) : *(const u32_a2*)(p + e);
    a = h2f(v & 0xffffu) + w;
    b = h2f(v >> 16) + w;
  }
  __device__ static void st2(T* p, int64_t e, float w, float a, float b) {
    uint32_t v = f2h(a - w) | (f2h(b - w) << 16);
    __builtin_nontemporal_store(v, (uint32_t*)(p + e));
  }
};

struct SBF16 {
  using T = uint16_t;
  static constexpr int BYTES = 2;
  __device__ static float ld(const T* p, int64_t e, float w) { return b2f(p[e]) + w; }
  __device__ static T enc(float f, float w) { return (T)f2b(f - w); }
  __device__ static T enc_dev(float d) { return (T)f2b(d); }
  __device__ static T rest(float) { return 0; }
  template <bool ALIGNED>
  __device__ static void ld2(const T* p, int64_t e, float w, float& a, float& b) {
    uint32_t v = ALIGNED ? *(const uint32_t*)(p + e) : *(const u32_a2*)(p + e);
    a = b2f(v & 0xffffu) + w;
    b = b2f(v >> 16) + w;
  }
  __device__ static void st2(T* p, int64_t e, float w, float a, float b) {
    uint32_t v = f2b(a - w) | (f2b(b - w) << 16);
    __builtin_nontemporal_store(v, (uint32_t*)(p + e));
  }
};

// ---------------------------------------------------------------------------------------------
// Kernel parameter block: the reference's 96-byte uniform (wgslLbm.ts:6-15) as derived by
// SolverGPU.writeParams (SolverGPU.ts:160-187), plus the fast-division constants.
// ---------------------------------------------------------------------------------------------
struct FastDiv {  // unsigned 32-bit division by an invariant d >= 1 (round-up multiplier method)
  uint32_t d, m, s;
  __host__ void init(uint32_t div) {
    d = div;
    uint32_t l = 0;
    while ((1ull << l) < div) l++;
    m = (uint32_t)((((1ull << l) - div) << 32) / div + 1);
    s = l;
  }
  __device__ uint32_t div(uint32_t x) const {
    uint32_t t = __umulhi(m, x);
    return s == 0 ? x : (t + ((x - t) >> 1)) >> (s - 1);
  }
};

struct KParams {
  int32_t nx, ny, nz, q;
  int64_t n;
  float tau0, csmag2, uin, rhoIn;
  int32_t ground, sideMode, collision, interp;
  float spongeStart, spongeNu, uGround, emaAlpha;
  float uRef, tauWall, spongeIn, contactH;
  int32_t refX, closure, pad0, pad1;
  FastDiv divx, divy;
};

struct KWheels {
  int32_t count, pad0, pad1, pad2;
  float4 w[32];  // w[2k] = (centre, 1/R), w[2k+1] = (axis, half width), SolverGPU.ts:325-335
};

// wgslLbm.ts:21-28
__device__ inline float3 wheel_vel(const KWheels& W, const KParams& P, uint32_t k, float3 p) {
  float4 a = W.w[2 * k], b = W.w[2 * k + 1];
  float3 d = make_float3(p.x - a.x, p.y - a.y, p.z - a.z);
  float3 cr = make_float3(b.y * d.z - b.z * d.y, b.z * d.x - b.x * d.z, b.x * d.y - b.y * d.x);
  float s = P.uGround * a.w;
  return make_float3(s * cr.x, s * cr.y, s * cr.z);
}

// wgslLbm.ts:32-43
__device__ inline float sdf_at(const uint32_t* sdf, int64_t id) {
  uint32_t s = sdf[id];
  return s == 0u ? -1.0f : __uint_as_float(~s);
}
__device__ inline float link_q(const uint32_t* sdf, const KParams& P, int64_t fluid, int64_t solid) {
  if (P.interp == 0) return 0.5f;
  float a = sdf_at(sdf, fluid), b = sdf_at(sdf, solid);
  if (a < 0.0f || b < 0.0f || a + b < 1e-6f) return 0.5f;
  return fminf(fmaxf(a / (a + b), 0.01f), 0.99f);
}

// ---------------------------------------------------------------------------------------------
// Collision: moments, non-equilibrium stress, Smagorinsky tau, regularized (or BGK) relaxation,
// outlet sponge and inlet absorbing layer (wgslLbm.ts:177-240). The update is written per
// opposite-direction pair (i, opp i): the even part (mass, stress) and the odd part (momentum)
// are formed once per pair, which halves the per-direction arithmetic. Algebraically identical
// to the reference's per-direction form.
// `nut_tau` (when >= 0) replaces the Smagorinsky tau by a closure-supplied relaxation time.
// ---------------------------------------------------------------------------------------------
template <class L, bool MAYBE_WALL>
__device__ __forceinline__ void collide(float (&f)[L::Q], float fxpos, bool wall, bool gwall, const KParams& P) {
  constexpr int Q = L::Q;
  float rho = 0.f, jx = 0.f, jy = 0.f, jz = 0.f;
  float sxx = 0.f, syy = 0.f, szz = 0.f, sxy = 0.f, sxz = 0.f, syz = 0.f;
#pragma unroll
  for (int i = 0; i < Q; i++) {
    const int cx = L::c(i, 0), cy = L::c(i, 1), cz = L::c(i, 2);
    rho += f[i];
    if (cx) jx += cx * f[i];
    if (cy) jy += cy * f[i];
    if (cz) jz += cz * f[i];
    if (cx) sxx += f[i];
    if (cy) syy += f[i];
    if (cz) szz += f[i];
    if (cx * cy) sxy += (cx * cy) * f[i];
    if (cx * cz) sxz += (cx * cz) * f[i];
    if (cy * cz) syz += (cy * cz) * f[i];
  }
  const float irho = 1.0f / rho;
  const float ux = jx * irho, uy = jy * irho, uz = (L::DIM == 3) ? jz * irho : 0.0f;
  const float usq = ux * ux + uy * uy + uz * uz;
  const float third = 1.0f / 3.0f;
  const float pxx = sxx - rho * (ux * ux + third);
  const float pyy = syy - rho * (uy * uy + third);
  const float pzz = (L::DIM == 3) ? szz - rho * (uz * uz + third) : 0.0f;
  const float pxy = sxy - rho * ux * uy;
  const float pxz = (L::DIM == 3) ? sxz - rho * ux * uz : 0.0f;
  const float pyz = (L::DIM == 3) ? syz - rho * uy * uz : 0.0f;
  const float piN = sqrtf(pxx * pxx + pyy * pyy + pzz * pzz + 2.0f * (pxy * pxy + pxz * pxz + pyz * pyz));
  const float sp = fminf(fmaxf((fxpos - P.spongeStart) / fmaxf((float)P.nx - P.spongeStart, 1.0f), 0.0f), 1.0f);
  float t0 = P.tau0 + 3.0f * P.spongeNu * sp * sp;
  if (MAYBE_WALL && gwall) t0 = fmaxf(t0, P.tauWall);
  const float tau = 0.5f * (t0 + sqrtf(t0 * t0 + 25.4558441f * P.csmag2 * piN * irho));
  const float om1 = 1.0f - 1.0f / tau;
  const float k1 = om1 * 4.5f;
  const bool useBGK = (MAYBE_WALL && wall) || P.collision == 1;
  const float si = fminf(fmaxf(1.0f - fxpos / fmaxf(P.spongeIn, 1.0f), 0.0f), 1.0f);
  const float sIn = P.spongeIn > 0.0f ? 0.08f * si * si : 0.0f;
  const float trP = (pxx + pyy + pzz) * third;
  const float U = P.uin;
  const float base_e = rho * (1.0f - 1.5f * usq);           // even part of feq / w
  const float base_r = base_e - k1 * trP;                    // regularized even part / w (without c-terms)
  float out[Q];
  if (!useBGK) {
    out[0] = L::w(0) * base_r;
#pragma unroll
    for (int i = 1; i < Q; i++) {
      const int o = opp<L>(i);
      if (o < i) continue;
      const int cx = L::c(i, 0), cy = L::c(i, 1), cz = L::c(i, 2);
      const float cu = cx * ux + cy * uy + cz * uz;
      const float qp = (float)(cx * cx) * pxx + (float)(cy * cy) * pyy + (float)(cz * cz) * pzz +
                       (float)(2 * cx * cy) * pxy + (float)(2 * cx * cz) * pxz + (float)(2 * cy * cz) * pyz;
      const float ev = L::w(i) * (base_r + 4.5f * rho * cu * cu + k1 * qp);
      const float od = (3.0f * L::w(i)) * rho * cu;
      out[i] = ev + od;
      out[o] = ev - od;
    }
  } else {
    out[0] = om1 * f[0] + (1.0f - om1) * (L::w(0) * base_e);
#pragma unroll
    for (int i = 1; i < Q; i++) {
      const int o = opp<L>(i);
      if (o < i) continue;
      const int cx = L::c(i, 0), cy = L::c(i, 1), cz = L::c(i, 2);
      const float cu = cx * ux + cy * uy + cz * uz;
      const float ev = L::w(i) * (base_e + 4.5f * rho * cu * cu);
      const float od = (3.0f * L::w(i)) * rho * cu;
      out[i] = (ev + od) + om1 * (f[i] - (ev + od));
      out[o] = (ev - od) + om1 * (f[o] - (ev - od));
    }
  }
  if (sIn > 0.0f) {
#pragma unroll
    for (int i = 0; i < Q; i++) {
      const int cx = L::c(i, 0);
      const float tgt = L::w(i) * (1.0f + 3.0f * cx * U + 4.5f * (cx * cx) * U * U - 1.5f * U * U);
      out[i] = out[i] * (1.0f - sIn) + tgt * sIn;
    }
  }
#pragma unroll
  for (int i = 0; i < Q; i++) f[i] = out[i];
}

// ---------------------------------------------------------------------------------------------
// General per-cell pull + boundaries (wgslLbm.ts:139-175, :262-291), exactly the reference's
// branch order. Returns false for a solid cell (out = w).
// ---------------------------------------------------------------------------------------------
template <class L, class S>
__device__ inline void update_general(const typename S::T* __restrict__ fin, const uint32_t* __restrict__ flags,
                                      const uint32_t* __restrict__ sdf, const uint8_t* __restrict__ kind,
                                      const KParams& P, const KWheels& W, int64_t idx, float (&f)[L::Q]) {
  constexpr int Q = L::Q, DIM = L::DIM;
  const int64_t n = P.n;
  const int NX = P.nx, NY = P.ny, NZ = P.nz;
  const uint32_t row = P.divx.div((uint32_t)idx);
  const int ix = (int)((uint32_t)idx - row * (uint32_t)NX);
  const uint32_t zz = P.divy.div(row);
  const int iy = (int)(row - zz * (uint32_t)NY);
  const int iz = (int)zz;
  const uint8_t kd = kind[idx];
  if (kd & KIND_SOLID) {
#pragma unroll
    for (int i = 0; i < Q; i++) f[i] = L::w(i);
    return;
  }
  auto cell = [&](int x, int y, int z) -> int64_t { return (int64_t)x + (int64_t)NX * ((int64_t)y + (int64_t)NY * z); };
  const float U = P.uin;
  bool wall = false, gwall = false;
  float mvSum = 0.f, mvW = 0.f;
  uint32_t wm = 0u;
  const float belt = (P.ground == 2 && !(iy == 0 && (kd & KIND_BELT_HOLD))) ? 1.0f : 0.0f;
  f[0] = S::ld(fin, idx, L::w(0));
#pragma unroll
  for (int i = 1; i < Q; i++) {
    const int cx = L::c(i, 0), cy = L::c(i, 1), cz = L::c(i, 2);
    const float w = L::w(i);
    int sx = ix - cx, sy = iy - cy, sz = DIM == 3 ? iz - cz : 0;
    if (cx < 0) sx = min(sx, NX - 1);
    const bool yOut = cy > 0 ? sy < 0 : (cy < 0 ? sy >= NY : false);
    const bool zOut = DIM == 3 ? (cz > 0 ? sz < 0 : (cz < 0 ? sz >= NZ : false)) : false;
    float v;
    if (cx > 0 && sx < 0) {
      const float cu = (float)cx * U;
      v = w * P.rhoIn * (1.0f + 3.0f * cu + 4.5f * cu * cu - 1.5f * U * U);
    } else if (cy > 0 && yOut && P.ground != 0) {
      wall = true; gwall = true;
      v = S::ld(fin, (int64_t)opp<L>(i) * n + idx, L::w(opp<L>(i)));
      if (cx != 0) v += belt * ((6.0f * w * cx) * P.uGround);
    } else {
      const int sxc = min(max(sx, 0), NX - 1);
      if (DIM == 3 && yOut && zOut) {
        v = S::ld(fin, (int64_t)mirYZ<L>(i) * n + cell(sxc, iy, iz), L::w(mirYZ<L>(i)));
      } else if (yOut) {
        v = S::ld(fin, (int64_t)mirY<L>(i) * n + cell(sxc, iy, DIM == 3 ? min(max(sz, 0), NZ - 1) : 0), L::w(mirY<L>(i)));
      } else if (DIM == 3 && zOut) {
        if (P.sideMode == 1) v = S::ld(fin, (int64_t)i * n + cell(sxc, sy, (sz + NZ) % NZ), w);
        else v = S::ld(fin, (int64_t)mirZ<L>(i) * n + cell(sxc, sy, iz), L::w(mirZ<L>(i)));
      } else {
        const int64_t sc = cell(sxc, sy, sz);
        const uint32_t fs = flags[sc];
        if (fs != 0u) {
          wall = true;
          const float fb = S::ld(fin, (int64_t)opp<L>(i) * n + idx, L::w(opp<L>(i)));
          float q = link_q(sdf, P, idx, sc);
          v = fb;
          if (q < 0.5f) {
            const int ax = ix + cx, ay = iy + cy, az = DIM == 3 ? iz + cz : 0;
            if (ax >= 0 && ay >= 0 && az >= 0 && ax < NX && ay < NY && az < NZ && flags[cell(ax, ay, az)] == 0u)
              v = 2.0f * q * fb + (1.0f - 2.0f * q) * S::ld(fin, (int64_t)opp<L>(i) * n + cell(ax, ay, az), L::w(opp<L>(i)));
            else
              q = 0.5f;
          } else if (q > 0.5f) {
            v = (0.5f / q) * fb + (1.0f - 0.5f / q) * S::ld(fin, (int64_t)i * n + idx, w);
          }
          if ((fs >> 8) != 0u) {
            const float3 p = make_float3(0.5f * (float)(ix + sxc) + 0.5f, 0.5f * (float)(iy + sy) + 0.5f,
                                         0.5f * (float)(iz + sz) + 0.5f);
            const float3 uw = wheel_vel(W, P, (fs >> 8) - 1u, p);
            const float mv = (6.0f * w) * ((float)cx * uw.x + (float)cy * uw.y + (float)cz * uw.z) * (q > 0.5f ? 0.5f / q : 1.0f);
            v += mv; mvSum += mv; mvW += w; wm |= 1u << i;
          }
        } else {
          v = S::ld(fin, (int64_t)i * n + sc, w);
        }
      }
    }
    f[i] = v;
  }
  if (mvW > 0.0f) {
    const float corr = mvSum / mvW;
#pragma unroll
    for (int i = 1; i < Q; i++)
      if (wm & (1u << i)) f[i] -= L::w(i) * corr;
  }
  collide<L, true>(f, (float)ix, wall, gwall, P);
}

}  // namespace wt
