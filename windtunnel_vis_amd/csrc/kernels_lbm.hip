// Lattice-Boltzmann kernels for gfx950: init, cell classification, the bulk fused
// pull-stream + collide kernel, the boundary kernel, and the macro / momentum-exchange reduction.
#include "lbm.hpp"
#include "launch.hpp"

namespace wt {

constexpr int BLOCK = 256;

// ---------------------------------------------------------------------------------------------
// init: equilibrium at (rho = 1, u = (U,0,0)) on fluid cells and at rest on solid cells, written
// to one buffer (the host calls it for both), wgslLbm.ts:296-341.
// ---------------------------------------------------------------------------------------------
template <class L, class S>
__global__ __launch_bounds__(BLOCK) void k_init(typename S::T* __restrict__ fout, const uint32_t* __restrict__ flags, KParams P) {
  const int64_t idx = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (idx >= P.n) return;
  const float U = flags[idx] != 0u ? 0.0f : P.uin;
#pragma unroll
  for (int i = 0; i < L::Q; i++) {
    const float cx = (float)L::c(i, 0);
    if constexpr (S::BYTES == 4)
      fout[(int64_t)i * P.n + idx] = L::w(i) * (1.0f + 3.0f * cx * U + 4.5f * cx * cx * U * U - 1.5f * U * U);
    else  // half-width formats store the deviation, computed directly (wgslLbm.ts:302)
      fout[(int64_t)i * P.n + idx] = S::enc_dev(L::w(i) * (3.0f * cx * U + 4.5f * cx * cx * U * U - 1.5f * U * U));
  }
}

// ---------------------------------------------------------------------------------------------
// classify: kind byte per cell (see lbm.hpp) and the list of x-pairs that need the boundary path.
// A fluid cell is "boundary" when any pull source x - c_i leaves the domain or is solid.
// ---------------------------------------------------------------------------------------------
template <class L>
__global__ __launch_bounds__(BLOCK) void k_classify(const uint32_t* __restrict__ flags, uint8_t* __restrict__ kind, KParams P) {
  const int64_t idx = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (idx >= P.n) return;
  const uint32_t row = P.divx.div((uint32_t)idx);
  const int ix = (int)((uint32_t)idx - row * (uint32_t)P.nx);
  const uint32_t zz = P.divy.div(row);
  const int iy = (int)(row - zz * (uint32_t)P.ny);
  const int iz = (int)zz;
  if (flags[idx] != 0u) { kind[idx] = KIND_SOLID; return; }
  bool b = ix == 0 || ix == P.nx - 1 || iy == 0 || iy == P.ny - 1;
  if (L::DIM == 3) b = b || iz == 0 || iz == P.nz - 1;
  if (!b) {
#pragma unroll
    for (int i = 1; i < L::Q; i++) {
      const int64_t s = idx - (L::c(i, 0) + (int64_t)P.nx * (L::c(i, 1) + (int64_t)P.ny * L::c(i, 2)));
      if (flags[s] != 0u) b = true;
    }
  }
  uint8_t k = b ? KIND_BOUNDARY : KIND_FLUID;
  if (iy == 0) {  // tyre contact patch scan, wgslLbm.ts:276-287 (contactH = 2, SolverGPU.ts:183)
    const int r = (int)P.contactH;
    bool hold = false;
    for (int dz = -r; dz <= r; dz++)
      for (int dx = -r; dx <= r; dx++)
        for (int dy = 0; dy <= 1; dy++) {
          const int qx = ix + dx, qz = iz + dz;
          if (qx >= 0 && qx < P.nx && qz >= 0 && qz < P.nz && dy < P.ny) {
            const uint32_t fq = flags[(int64_t)qx + (int64_t)P.nx * ((int64_t)dy + (int64_t)P.ny * qz)];
            if (fq != 0u && (fq >> 8) == 0u) hold = true;
          }
        }
    if (hold) k |= KIND_BELT_HOLD;
  }
  kind[idx] = k;
}

// Pairs (2p, 2p+1) with a boundary cell go to the list; counts[0] = list length,
// counts[1] = fast pairs, counts[2] = all-solid pairs.
__global__ __launch_bounds__(BLOCK) void k_collect(const uint16_t* __restrict__ kind2, int64_t npairs, uint32_t* __restrict__ list,
                                                   unsigned long long* __restrict__ counts) {
  const int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  bool bnd = false, fast = false, sol = false;
  if (p < npairs) {
    const uint16_t k = kind2[p];
    fast = (k == 0);
    sol = (k == 0x0101u);
    bnd = !fast && !sol;
  }
  // wave-aggregated append
  const unsigned long long mb = __ballot(bnd);
  const unsigned long long mf = __ballot(fast);
  const unsigned long long ms = __ballot(sol);
  const int lane = threadIdx.x & 63;
  uint32_t base = 0;
  if (lane == 0) {
    if (mb) base = (uint32_t)atomicAdd(&counts[0], (unsigned long long)__popcll(mb));
    if (mf) atomicAdd(&counts[1], (unsigned long long)__popcll(mf));
    if (ms) atomicAdd(&counts[2], (unsigned long long)__popcll(ms));
  }
  base = __shfl(base, 0);
  if (bnd) {
    const int rank = __popcll(mb & ((1ull << lane) - 1ull));
    list[base + rank] = (uint32_t)p;
  }
}

// ---------------------------------------------------------------------------------------------
// Bulk kernel: one thread per x-pair of cells. Pairs whose two cells are interior fluid take the
// fast path (plain pulls from in-domain fluid neighbours, then collision); all-solid pairs write
// the rest state; pairs with a boundary cell are left to the boundary kernel.
// HBM traffic per fast cell: Q loads + Q stores of S::BYTES (the algorithmic 2 Q s), plus one
// kind byte.
// ---------------------------------------------------------------------------------------------
template <class L, class S>
__global__ __launch_bounds__(BLOCK) void k_bulk(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                                const uint16_t* __restrict__ kind2, KParams P) {
  const int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const int64_t n = P.n;
  if (p >= (n >> 1)) return;
  const uint16_t k = kind2[p];
  const int64_t c0 = 2 * p;
  if (k == 0x0101u) {
#pragma unroll
    for (int i = 0; i < L::Q; i++) S::st2(fout, (int64_t)i * n + c0, L::w(i), L::w(i), L::w(i));
    return;
  }
  if (k != 0) return;
  float fa[L::Q], fb[L::Q];
#pragma unroll
  for (int i = 0; i < L::Q; i++) {
    const int64_t off = L::c(i, 0) + (int64_t)P.nx * (L::c(i, 1) + (int64_t)P.ny * L::c(i, 2));
    const int64_t e = (int64_t)i * n + c0 - off;
    if (L::c(i, 0) == 0) S::template ld2<true>(fin, e, L::w(i), fa[i], fb[i]);
    else S::template ld2<false>(fin, e, L::w(i), fa[i], fb[i]);
  }
  const uint32_t row = P.divx.div((uint32_t)c0);
  const float x0 = (float)((uint32_t)c0 - row * (uint32_t)P.nx);
  collide<L, false>(fa, x0, false, false, P);
  collide<L, false>(fb, x0 + 1.0f, false, false, P);
#pragma unroll
  for (int i = 0; i < L::Q; i++) S::st2(fout, (int64_t)i * n + c0, L::w(i), fa[i], fb[i]);
}

// Boundary kernel: one thread per listed x-pair, the reference's full branch logic per cell.
template <class L, class S>
__global__ __launch_bounds__(BLOCK) void k_boundary(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                                    const uint32_t* __restrict__ flags, const uint32_t* __restrict__ sdf,
                                                    const uint8_t* __restrict__ kind, const uint32_t* __restrict__ list,
                                                    int64_t count, KParams P, KWheels W) {
  const int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (t >= count) return;
  const int64_t c0 = 2 * (int64_t)list[t];
  float fa[L::Q], fb[L::Q];
  update_general<L, S>(fin, flags, sdf, kind, P, W, c0, fa);
  if (c0 + 1 < P.n) {
    update_general<L, S>(fin, flags, sdf, kind, P, W, c0 + 1, fb);
#pragma unroll
    for (int i = 0; i < L::Q; i++) S::st2(fout, (int64_t)i * P.n + c0, L::w(i), fa[i], fb[i]);
  } else {  // odd cell count (FP32 storage only): the last cell has no partner
#pragma unroll
    for (int i = 0; i < L::Q; i++) fout[(int64_t)i * P.n + c0] = S::enc(fa[i], L::w(i));
  }
}

// ---------------------------------------------------------------------------------------------
// Macro pass (wgslLbm.ts:353-489): rho, u -> vel; EMA statistics; momentum exchange on links into
// the body with the Bouzidi / wheel terms; NaN flag; reference-plane density; solid density
// extrapolation. Block partials are reduced in FP64 (force sums) and max (u^2).
// partials layout per block: [Fx, Fy, Fz, max u^2, sum rho_ref, n_ref, nan, solid] (double).
// ---------------------------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ inline float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

template <class L, class S>
__global__ __launch_bounds__(BLOCK) void k_macro(const typename S::T* __restrict__ fin, const uint32_t* __restrict__ flags,
                                                 const uint32_t* __restrict__ sdf, const uint8_t* __restrict__ kind,
                                                 ushort4* __restrict__ vel, float4* __restrict__ meanA, float* __restrict__ meanB,
                                                 double* __restrict__ partials, KParams P, KWheels W) {
  constexpr int Q = L::Q, DIM = L::DIM;
  const int64_t idx = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const int64_t n = P.n;
  double F0 = 0, F1 = 0, F2 = 0, B0 = 0, B1 = 0, B2 = 0, B3 = 0;
  float amax = 0.0f;
  if (idx < n) {
    const uint32_t row = P.divx.div((uint32_t)idx);
    const int ix = (int)((uint32_t)idx - row * (uint32_t)P.nx);
    const uint32_t zz = P.divy.div(row);
    const int iy = (int)(row - zz * (uint32_t)P.ny);
    const int iz = (int)zz;
    const int NX = P.nx, NY = P.ny, NZ = P.nz;
    auto safe = [&](int x, int y, int z) -> int64_t {
      if (x < 0 || y < 0 || z < 0 || x >= NX || y >= NY || z >= NZ) return -1;
      return (int64_t)x + (int64_t)NX * ((int64_t)y + (int64_t)NY * z);
    };
    float4 out = make_float4(0.f, 0.f, 0.f, 1.f);
    const uint8_t kd = kind[idx];
    const float a = P.emaAlpha;
    if (!(kd & KIND_SOLID)) {
      float f[Q];
#pragma unroll
      for (int i = 0; i < Q; i++) f[i] = S::ld(fin, (int64_t)i * n + idx, L::w(i));
      float rho = 0.f, jx = 0.f, jy = 0.f, jz = 0.f;
#pragma unroll
      for (int i = 0; i < Q; i++) {
        rho += f[i];
        if (L::c(i, 0)) jx += L::c(i, 0) * f[i];
        if (L::c(i, 1)) jy += L::c(i, 1) * f[i];
        if (L::c(i, 2)) jz += L::c(i, 2) * f[i];
      }
      const float ux = jx / rho, uy = jy / rho, uz = DIM == 3 ? jz / rho : 0.0f;
      out = make_float4(ux, uy, uz, rho);
      float Fx = 0.f, Fy = 0.f, Fz = 0.f, mvSum = 0.f, mvW = 0.f, Cx = 0.f, Cy = 0.f, Cz = 0.f;
      if (kd & KIND_BOUNDARY) {
#pragma unroll
        for (int i = 1; i < Q; i++) {
          const int cx = L::c(i, 0), cy = L::c(i, 1), cz = L::c(i, 2);
          const int64_t s = safe(ix + cx, iy + cy, iz + cz);
          if (s < 0 || flags[s] == 0u) continue;
          float q = link_q(sdf, P, idx, s);
          float fr = f[i];
          if (q < 0.5f) {
            const int64_t bk = safe(ix - cx, iy - cy, iz - cz);
            if (bk >= 0 && flags[bk] == 0u) fr = 2.0f * q * f[i] + (1.0f - 2.0f * q) * S::ld(fin, (int64_t)i * n + bk, L::w(i));
            else q = 0.5f;
          } else if (q > 0.5f) {
            fr = (0.5f / q) * f[i] + (1.0f - 0.5f / q) * f[opp<L>(i)];
          }
          const uint32_t tg = flags[s] >> 8;
          if (tg != 0u) {
            const float3 pw = make_float3((float)ix + (0.5f + 0.5f * cx), (float)iy + (0.5f + 0.5f * cy), (float)iz + (0.5f + 0.5f * cz));
            const float3 uw = wheel_vel(W, P, tg - 1u, pw);
            const float mv = -(6.0f * L::w(i)) * (q > 0.5f ? 0.5f / q : 1.0f) * ((float)cx * uw.x + (float)cy * uw.y + (float)cz * uw.z);
            fr += mv; mvSum += mv; mvW += L::w(i);
            Cx += cx * L::w(i); Cy += cy * L::w(i); Cz += cz * L::w(i);
          }
          const float m = f[i] + fr - 2.0f * L::w(i);
          Fx += cx * m; Fy += cy * m; Fz += cz * m;
        }
        if (mvW > 0.0f) {
          const float r = mvSum / mvW;
          Fx -= Cx * r; Fy -= Cy * r; Fz -= Cz * r;
        }
      }
      const float uu = ux * ux + uy * uy + uz * uz;
      const bool bad = !(uu < 1.0e6f) || !(rho > 0.0f);
      F0 = Fx; F1 = Fy; F2 = Fz;
      amax = bad ? 1.0e6f : uu;
      if (ix == P.refX) { B0 = rho; B1 = 1.0; }
      B2 = bad ? 1.0 : 0.0;
      const float4 m = meanA[idx];
      meanA[idx] = make_float4(m.x * (1.0f - a) + out.x * a, m.y * (1.0f - a) + out.y * a, m.z * (1.0f - a) + out.z * a,
                               m.w * (1.0f - a) + out.w * a);
      meanB[idx] = meanB[idx] * (1.0f - a) + uu * a;
    } else {
      float rs = 0.f, rn = 0.f;
      const int nb[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
#pragma unroll
      for (int k = 0; k < (DIM == 3 ? 6 : 4); k++) {
        const int64_t s = safe(ix + nb[k][0], iy + nb[k][1], iz + nb[k][2]);
        if (s >= 0 && flags[s] == 0u) {
          float r = 0.f;
#pragma unroll
          for (int i = 0; i < Q; i++) r += S::ld(fin, (int64_t)i * n + s, L::w(i));
          rs += r; rn += 1.0f;
        }
      }
      if (rn > 0.0f) out.w = rs / rn;
      const float4 m = meanA[idx];
      meanA[idx] = make_float4(m.x * (1.0f - a) + out.x * a, m.y * (1.0f - a) + out.y * a, m.z * (1.0f - a) + out.z * a,
                               m.w * (1.0f - a) + out.w * a);
      meanB[idx] = 0.0f;
      B3 = 1.0;
    }
    vel[idx] = make_ushort4((unsigned short)f2h(out.x), (unsigned short)f2h(out.y), (unsigned short)f2h(out.z),
                            (unsigned short)f2h(out.w));
  }
  // block reduction: waves via cross-lane shuffles, then LDS across the 4 waves
  F0 = wave_sum(F0); F1 = wave_sum(F1); F2 = wave_sum(F2);
  B0 = wave_sum(B0); B1 = wave_sum(B1); B2 = wave_sum(B2); B3 = wave_sum(B3);
  amax = wave_max(amax);
  __shared__ double sh[BLOCK / 64][8];
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
    sh[wid][0] = F0; sh[wid][1] = F1; sh[wid][2] = F2; sh[wid][3] = (double)amax;
    sh[wid][4] = B0; sh[wid][5] = B1; sh[wid][6] = B2; sh[wid][7] = B3;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    const int k = threadIdx.x;
    double v = sh[0][k];
    for (int w = 1; w < BLOCK / 64; w++) v = (k == 3) ? fmax(v, sh[w][k]) : v + sh[w][k];
    partials[(int64_t)blockIdx.x * 8 + k] = v;
  }
}

// Final reduction of the block partials (wgslLbm.ts:491-521) into result[8], FP64.
__global__ __launch_bounds__(1024) void k_reduce(const double* __restrict__ partials, int64_t count, double* __restrict__ result) {
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t b = threadIdx.x; b < count; b += 1024) {
#pragma unroll
    for (int k = 0; k < 8; k++) acc[k] = (k == 3) ? fmax(acc[k], partials[b * 8 + k]) : acc[k] + partials[b * 8 + k];
  }
  __shared__ double sh[16][8];
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    double v = acc[k];
    for (int o = 32; o > 0; o >>= 1) {
      const double u = __shfl_xor(v, o);
      v = (k == 3) ? fmax(v, u) : v + u;
    }
    if (lane == 0) sh[wid][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    const int k = threadIdx.x;
    double v = sh[0][k];
    for (int w = 1; w < 16; w++) v = (k == 3) ? fmax(v, sh[w][k]) : v + sh[w][k];
    result[k] = v;
  }
}

// Convert between device storage and full-value host floats.
template <class L, class S>
__global__ __launch_bounds__(BLOCK) void k_unpack(const typename S::T* __restrict__ f, float* __restrict__ out, KParams P) {
  const int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (e >= (int64_t)L::Q * P.n) return;
  const int i = (int)(e / P.n);
  float w = 0.0f;
#pragma unroll
  for (int k = 0; k < L::Q; k++) if (k == i) w = L::w(k);
  out[e] = S::ld(f, e, w);
}
template <class L, class S>
__global__ __launch_bounds__(BLOCK) void k_pack(const float* __restrict__ in, typename S::T* __restrict__ f, KParams P) {
  const int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (e >= (int64_t)L::Q * P.n) return;
  const int i = (int)(e / P.n);
  float w = 0.0f;
#pragma unroll
  for (int k = 0; k < L::Q; k++) if (k == i) w = L::w(k);
  f[e] = S::enc(in[e], w);
}

__global__ __launch_bounds__(BLOCK) void k_clear_stats(float4* __restrict__ meanA, float* __restrict__ meanB, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  meanA[i] = make_float4(0.f, 0.f, 0.f, 1.f);
  meanB[i] = 0.0f;
}

// ---------------------------------------------------------------------------------------------
// Host-side launchers (templated dispatch on lattice and storage).
// ---------------------------------------------------------------------------------------------
static inline unsigned blocks_for(int64_t n, int b = BLOCK) { return (unsigned)((n + b - 1) / b); }

template <class F>
static void dispatch(int q, int prec, F&& fn) {
  if (q == 9) {
    if (prec == 0) fn(D2Q9{}, SF32{});
    else if (prec == 1) fn(D2Q9{}, SF16{});
    else fn(D2Q9{}, SBF16{});
  } else {
    if (prec == 0) fn(D3Q19{}, SF32{});
    else if (prec == 1) fn(D3Q19{}, SF16{});
    else fn(D3Q19{}, SBF16{});
  }
}

void launch_init(const LaunchCtx& c, void* f) {
  dispatch(c.P.q, c.prec, [&](auto l, auto s) {
    using L = decltype(l); using S = decltype(s);
    hipLaunchKernelGGL((k_init<L, S>), dim3(blocks_for(c.P.n)), dim3(BLOCK), 0, c.stream, (typename S::T*)f, c.flags, c.P);
  });
}

void launch_classify(const LaunchCtx& c, uint32_t* list, unsigned long long* counts) {
  if (c.P.q == 9) hipLaunchKernelGGL((k_classify<D2Q9>), dim3(blocks_for(c.P.n)), dim3(BLOCK), 0, c.stream, c.flags, c.kind, c.P);
  else hipLaunchKernelGGL((k_classify<D3Q19>), dim3(blocks_for(c.P.n)), dim3(BLOCK), 0, c.stream, c.flags, c.kind, c.P);
  const int64_t npairs = (c.P.n + 1) / 2;  // an odd last cell forms a pair with the zero padding
  hipLaunchKernelGGL(k_collect, dim3(blocks_for(npairs)), dim3(BLOCK), 0, c.stream, (const uint16_t*)c.kind, npairs, list, counts);
}

void launch_bulk(const LaunchCtx& c, const void* fin, void* fout) {
  dispatch(c.P.q, c.prec, [&](auto l, auto s) {
    using L = decltype(l); using S = decltype(s);
    hipLaunchKernelGGL((k_bulk<L, S>), dim3(blocks_for(c.P.n / 2)), dim3(BLOCK), 0, c.stream,
                       (const typename S::T*)fin, (typename S::T*)fout, (const uint16_t*)c.kind, c.P);
  });
}

void launch_boundary(const LaunchCtx& c, const void* fin, void* fout, const uint32_t* list, int64_t count) {
  if (count <= 0) return;
  dispatch(c.P.q, c.prec, [&](auto l, auto s) {
    using L = decltype(l); using S = decltype(s);
    hipLaunchKernelGGL((k_boundary<L, S>), dim3(blocks_for(count)), dim3(BLOCK), 0, c.stream, (const typename S::T*)fin,
                       (typename S::T*)fout, c.flags, c.sdf, c.kind, list, count, c.P, *c.W);
  });
}

int64_t macro_blocks(int64_t n) { return (int64_t)blocks_for(n); }

void launch_macro(const LaunchCtx& c, const void* fin, void* vel, void* meanA, float* meanB, double* partials, double* result) {
  dispatch(c.P.q, c.prec, [&](auto l, auto s) {
    using L = decltype(l); using S = decltype(s);
    hipLaunchKernelGGL((k_macro<L, S>), dim3(blocks_for(c.P.n)), dim3(BLOCK), 0, c.stream, (const typename S::T*)fin, c.flags,
                       c.sdf, c.kind, (ushort4*)vel, (float4*)meanA, meanB, partials, c.P, *c.W);
  });
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, c.stream, partials, (int64_t)blocks_for(c.P.n), result);
}

void launch_unpack(const LaunchCtx& c, const void* f, float* out) {
  dispatch(c.P.q, c.prec, [&](auto l, auto s) {
    using L = decltype(l); using S = decltype(s);
    hipLaunchKernelGGL((k_unpack<L, S>), dim3(blocks_for((int64_t)L::Q * c.P.n)), dim3(BLOCK), 0, c.stream,
                       (const typename S::T*)f, out, c.P);
  });
}

void launch_pack(const LaunchCtx& c, const float* in, void* f) {
  dispatch(c.P.q, c.prec, [&](auto l, auto s) {
    using L = decltype(l); using S = decltype(s);
    hipLaunchKernelGGL((k_pack<L, S>), dim3(blocks_for((int64_t)L::Q * c.P.n)), dim3(BLOCK), 0, c.stream, in,
                       (typename S::T*)f, c.P);
  });
}

void launch_clear_stats(const LaunchCtx& c, void* meanA, float* meanB) {
  hipLaunchKernelGGL(k_clear_stats, dim3(blocks_for(c.P.n)), dim3(BLOCK), 0, c.stream, (float4*)meanA, meanB, c.P.n);
}

}  // namespace wt
