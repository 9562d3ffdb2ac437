// Lattice-Boltzmann kernels for gfx950: init, cell classification, the bulk fused
// pull-stream + collide kernel, the wall kernel, and the macro / momentum-exchange reduction.
#include <algorithm>

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
// Classification. A fluid cell is WALL when the reference's plain pull branch (wgslLbm.ts:172) of
// some direction reads a solid cell; inlet, ground, mirror and periodic pulls never test solids.
// ---------------------------------------------------------------------------------------------
template <class L>
__device__ inline uint32_t solid_mask(const uint32_t* __restrict__ flags, int ix, int iy, int iz, const KParams& P) {
  uint32_t m = 0u;
#pragma unroll
  for (int i = 1; i < L::Q; i++) {
    const int cx = L::c(i, 0), cy = L::c(i, 1), cz = L::c(i, 2);
    int sx = ix - cx;
    const int sy = iy - cy, sz = L::DIM == 3 ? iz - cz : 0;
    if (cx < 0) sx = min(sx, P.nx - 1);
    if (cx > 0 && sx < 0) continue;
    if (sy < 0 || sy >= P.ny || sz < 0 || sz >= P.nz) continue;
    if (flags[(int64_t)sx + (int64_t)P.nx * ((int64_t)sy + (int64_t)P.ny * sz)] != 0u) m |= 1u << i;
  }
  return m;
}

__device__ inline void coords(int64_t idx, const KParams& P, int& ix, int& iy, int& iz) {
  const uint32_t row = P.divx.div((uint32_t)idx);
  ix = (int)((uint32_t)idx - row * (uint32_t)P.nx);
  const uint32_t zz = P.divy.div(row);
  iy = (int)(row - zz * (uint32_t)P.ny);
  iz = (int)zz;
}

template <class L>
__global__ __launch_bounds__(BLOCK) void k_classify(const uint32_t* __restrict__ flags, uint8_t* __restrict__ kind, KParams P) {
  const int64_t idx = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (idx >= P.n) return;
  int ix, iy, iz;
  coords(idx, P, ix, iy, iz);
  if (flags[idx] != 0u) { kind[idx] = KIND_SOLID; return; }
  uint8_t k = solid_mask<L>(flags, ix, iy, iz, P) ? KIND_WALL : KIND_FLUID;
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

// Groups of CPT cells along x (the bulk kernel's unit) that contain a WALL cell, or mix solid and
// fluid, go to the wall kernel; their cells are listed with their solid masks.
// counts[0] = listed cells, counts[1] = bulk fluid cells, counts[2] = bulk solid cells.
// Per-link data of a wall cell, fixed by the geometry: the Bouzidi fraction q of every solid link
// (with the reference's fallback to q = 1/2 when the second fluid node is missing folded in,
// wgslLbm.ts:172) and the wheel tag of the solid cell it points into.
template <class L>
__device__ WallRec wall_record(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ sdf, int64_t c, int ix, int iy,
                               int iz, uint32_t m, KParams P) {
  P.interp = 1;  // the record holds the interpolated q; halfway walls select 1/2 at step time
  WallRec R;
#pragma unroll
  for (int i = 0; i < 20; i++) { R.q[i < 19 ? i : 18] = 0.5f; R.tag[i] = 0; }
#pragma unroll
  for (int i = 1; i < L::Q; i++) {
    if (!(m & (1u << i))) continue;
    const int cx = L::c(i, 0), cy = L::c(i, 1), cz = L::c(i, 2);
    const int sx = cx < 0 ? min(ix - cx, P.nx - 1) : ix - cx;
    const int64_t sc = (int64_t)sx + (int64_t)P.nx * ((int64_t)(iy - cy) + (int64_t)P.ny * (L::DIM == 3 ? iz - cz : 0));
    float q = link_q(sdf, P, c, sc);
    if (q < 0.5f) {
      const int ax = ix + cx, ay = iy + cy, az = L::DIM == 3 ? iz + cz : 0;
      const bool ok = ax >= 0 && ay >= 0 && az >= 0 && ax < P.nx && ay < P.ny && az < P.nz &&
                      flags[(int64_t)ax + (int64_t)P.nx * ((int64_t)ay + (int64_t)P.ny * az)] == 0u;
      if (!ok) q = 0.5f;
    }
    R.q[i] = q;
    R.tag[i] = (uint8_t)(flags[sc] >> 8);
  }
  return R;
}

template <class L, int CPT>
__global__ __launch_bounds__(BLOCK) void k_collect(const uint8_t* __restrict__ kind, const uint32_t* __restrict__ flags,
                                                   uint32_t* __restrict__ list, uint32_t* __restrict__ masks,
                                                   unsigned long long* __restrict__ counts, KParams P) {
  const int64_t g = (int64_t)blockIdx.x * BLOCK + threadIdx.x;  // group index
  const int64_t ngroups = (P.n + CPT - 1) / CPT;
  int nl = 0, nf = 0, ns = 0;
  const int64_t c0 = g * CPT;
  if (g < ngroups) {
    bool anyw = false, anys = false, alls = true;
#pragma unroll
    for (int j = 0; j < CPT; j++) {
      const int kj = (c0 + j < P.n) ? kind[c0 + j] : KIND_SOLID;
      anyw |= (kj & KIND_WALL) != 0;
      anys |= (kj & KIND_SOLID) != 0;
      alls &= (kj & KIND_SOLID) != 0;
    }
    // odd cell counts: a group that runs past the end is always listed
    const bool tail = c0 + CPT > P.n;
    if (anyw || (anys && !alls) || tail) nl = (int)min<int64_t>(CPT, P.n - c0);
    else if (alls) ns = CPT;
    else nf = CPT;
  }
  // wave-aggregated append
  const int lane = threadIdx.x & 63;
  int incl = nl;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  const int total = __shfl(incl, 63);
  unsigned long long base = 0;
  int tf = nf, ts = ns;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) { tf += __shfl_xor(tf, o); ts += __shfl_xor(ts, o); }
  if (lane == 0) {
    if (total) base = atomicAdd(&counts[0], (unsigned long long)total);
    if (tf) atomicAdd(&counts[1], (unsigned long long)tf);
    if (ts) atomicAdd(&counts[2], (unsigned long long)ts);
  }
  base = __shfl(base, 0);
  if (nl) {
    const uint32_t pos = (uint32_t)(base + incl - nl);
    for (int j = 0; j < nl; j++) {
      const int64_t c = c0 + j;
      int ix, iy, iz;
      coords(c, P, ix, iy, iz);
      list[pos + j] = (uint32_t)c;
      masks[pos + j] = (kind[c] & KIND_SOLID) ? 0u : solid_mask<L>(flags, ix, iy, iz, P);
    }
  }
}

// Per-link records of the listed wall cells; rerun whenever flags or distances change.
template <class L>
__global__ __launch_bounds__(BLOCK) void k_wallprep(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ sdf,
                                                    const uint32_t* __restrict__ list, const uint32_t* __restrict__ masks,
                                                    WallRec* __restrict__ recs, int64_t count, KParams P) {
  const int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (t >= count) return;
  const int64_t c = list[t];
  int ix, iy, iz;
  coords(c, P, ix, iy, iz);
  recs[t] = wall_record<L>(flags, sdf, c, ix, iy, iz, masks[t], P);
}

// ---------------------------------------------------------------------------------------------
// Bulk kernel. Block = 4 waves; each wave owns a chunk of 64 * CPT cells of one x-row, so the row
// (y, z) — and with it every face condition (ground, mirrors, periodic z, y-z edges) — is uniform
// across the wave and is resolved with scalar arithmetic. For each direction the kernel picks the
// source direction and source row once per wave, then every lane loads CPT consecutive cells
// (x shifted by -c_x; unaligned dword loads for FP16/BF16 pairs). Inlet (x = 0) and outlet clamp
// (x = nx - 1) are per-lane fix-ups. Groups holding a WALL cell are left to the wall kernel.
// HBM traffic per cell: Q loads + Q stores of S::BYTES (the algorithmic 2 Q s) + 1 kind byte.
// ---------------------------------------------------------------------------------------------
template <class L, class S, int CPT>
__device__ __forceinline__ void bulk_cells(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                           const uint8_t* __restrict__ kind, const KParams& P, uint32_t row, int x0) {
  using T = typename S::T;
  constexpr int Q = L::Q;
  constexpr bool D3 = L::DIM == 3;
  const int NX = P.nx, NY = P.ny, NZ = P.nz;
  if (x0 >= NX) return;
  const int64_t n = P.n;
  const int iz = (int)(row / (uint32_t)NY);
  const int iy = (int)row - iz * NY;
  const int64_t rowoff = (int64_t)row * NX;
  const int64_t c0 = rowoff + x0;
  // the kind byte(s) are loaded together with the populations: every source address below is in
  // bounds whatever the kind, so the loads need not wait for the skip test
  uint32_t k;
  if constexpr (CPT == 2) k = *(const uint16_t*)(kind + c0);
  else k = kind[c0];
  constexpr uint32_t SOLID_ALL = CPT == 2 ? 0x0101u : 0x01u;
  constexpr uint32_t SKIP = CPT == 2 ? 0x0303u : 0x03u;
  const bool yLo = iy == 0, yHi = iy == NY - 1;
  const bool zLo = D3 && iz == 0, zHi = D3 && iz == NZ - 1;
  const bool gwall = yLo && P.ground != 0;
  // ---- 1. per direction: source direction, row and x shift (wave-uniform, branch-free) ----
  const T* src[Q];
  int xlim[Q];  // largest x offset a lane may use (outlet clamp for CPT = 1)
#pragma unroll
  for (int i = 0; i < Q; i++) {
    const int cx = L::c(i, 0), cy = L::c(i, 1), cz = L::c(i, 2);
    const bool yOut = (cy > 0 && yLo) || (cy < 0 && yHi);
    const bool zOut = (cz > 0 && zLo) || (cz < 0 && zHi);
    const bool grnd = cy > 0 && yOut && P.ground != 0;
    const bool yz = D3 && yOut && zOut;
    const bool zper = P.sideMode == 1;
    // source direction: opp (ground), mirrorYZ, mirrorY, i / mirrorZ (z faces), i
    const int d = grnd ? opp<L>(i) : yz ? mirYZ<L>(i) : yOut ? mirY<L>(i) : (zOut && !zper) ? mirZ<L>(i) : i;
    const int sy = (grnd || yOut) ? iy : iy - cy;
    const int szp = D3 ? ((iz - cz + NZ) % NZ) : 0;
    const int sz = (grnd || yz) ? iz : yOut ? (D3 ? min(max(iz - cz, 0), NZ - 1) : 0) : zOut ? (zper ? szp : iz) : iz - cz;
    const int xs = grnd ? 0 : -cx;
    src[i] = fin + (int64_t)d * n + ((int64_t)sy + (int64_t)NY * sz) * NX + xs;
    xlim[i] = (grnd || cx >= 0) ? NX - 1 : NX - 2;
  }
  // ---- 2. all loads in flight together (scalar base + 32-bit lane byte offset) ----
  const uint32_t lane_off = (uint32_t)x0 * (uint32_t)sizeof(T);
  float fa[Q], fb[Q];
  if constexpr (CPT == 2) {
    typename S::Word raw[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) raw[i] = S::ld_word((const T*)((const char*)src[i] + lane_off));
#pragma unroll
    for (int i = 0; i < Q; i++) S::unpack_word(raw[i], L::w(i), fa[i], fb[i]);
  } else {
    T raw[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) raw[i] = src[i][min(x0, xlim[i])];
#pragma unroll
    for (int i = 0; i < Q; i++) {
      fa[i] = S::dev_of(raw[i], L::w(i));
      fb[i] = 0.0f;
    }
  }
  if (k == SOLID_ALL) {
#pragma unroll
    for (int i = 0; i < Q; i++) {
      T* o = fout + (int64_t)i * n + c0;
      if constexpr (CPT == 2) S::st2_dev(o, L::w(i), 0.0f, 0.0f);
      else S::st_dev(o, 0, L::w(i), 0.0f);
    }
    return;
  }
  if (k & SKIP) return;
  // ---- 3. per-lane fix-ups: outlet clamp, moving belt, inlet ----
  const float beltA = (P.ground == 2 && !(k & KIND_BELT_HOLD)) ? 1.0f : 0.0f;
  const float beltB = (P.ground == 2 && !((k >> 8) & KIND_BELT_HOLD)) ? 1.0f : 0.0f;
  const float U = P.uin;
#pragma unroll
  for (int i = 1; i < Q; i++) {
    const int cx = L::c(i, 0), cy = L::c(i, 1);
    const bool grnd = cy > 0 && yLo && P.ground != 0;
    if constexpr (CPT == 2) {
      if (cx < 0 && !grnd && x0 + 2 == NX) fb[i] = fa[i];  // outlet: x = nx clamps to nx - 1
    }
    if (grnd && cx != 0) {  // moving belt, wgslLbm.ts:159-162
      const float m = (6.0f * L::w(i) * cx) * P.uGround;
      fa[i] += beltA * m;
      fb[i] += beltB * m;
    }
    if (cx > 0 && x0 == 0) fa[i] = inlet_dev<L>(i, U);  // inlet, wgslLbm.ts:146 (checked first there)
  }
  collide<L>(fa, (float)x0, gwall, gwall, P);
  if constexpr (CPT == 2) collide<L>(fb, (float)(x0 + 1), gwall, gwall, P);
#pragma unroll
  for (int i = 0; i < Q; i++) {
    T* o = (T*)((char*)(fout + uniform_i64((int64_t)i * n + rowoff)) + lane_off);
    if constexpr (CPT == 2) S::st2_dev(o, L::w(i), fa[i], fb[i]);
    else S::st_dev(o, 0, L::w(i), fa[i]);
  }
}

// Wall cells: one lane per listed cell (cells next to the body), the reference's full branch
// logic with precomputed solid masks and link records.
template <class L, class S>
__device__ __forceinline__ void wall_cell(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                          const uint8_t* __restrict__ kind, const StepArgs& A, const KParams& P,
                                          const KWheels& W, int64_t t) {
  if (t >= A.nwall) return;
  const int64_t c = A.wlist[t];
  int ix, iy, iz;
  coords(c, P, ix, iy, iz);
  float f[L::Q];
  update_wall<L, S>(fin, kind[c], A.wmask[t], A.wrec[t], P, W, c, ix, iy, iz, f);
#pragma unroll
  for (int i = 0; i < L::Q; i++) S::st_dev(fout, (int64_t)i * P.n + c, L::w(i), f[i]);
}

// One lattice step in one launch. The first A.nwb blocks update the wall cells (long dependency
// chains, few lanes: started first so their latency hides under the bulk work); every other block
// is 4 waves of the bulk update, waves enumerated (row, chunk) with the chunk fastest so a block
// covers contiguous memory. With A.xcd set, bulk blocks are remapped so that each XCD (blocks are
// dealt round-robin to the 8 XCDs) walks one contiguous eighth of the grid, keeping the cache
// lines shared by neighbouring chunks inside one L2.
template <class L, class S, int CPT>
__global__ __launch_bounds__(BLOCK) void k_step(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                                const uint8_t* __restrict__ kind, KParams P, KWheels W, StepArgs A) {
  uint32_t b = blockIdx.x;
  if (b < A.nwb) {
    wall_cell<L, S>(fin, fout, kind, A, P, W, (int64_t)b * BLOCK + threadIdx.x);
    return;
  }
  b -= A.nwb;
  if (A.xcd) {
    const uint32_t per = (A.nbb + 7u) / 8u;
    const uint32_t lb = (b % 8u) * per + b / 8u;
    if (lb >= A.nbb) return;
    b = lb;
  }
  const uint32_t wave = b * 4u + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (wave >= A.nwaves) return;
  const uint32_t row = wave / A.wpr;
  const uint32_t chunk = wave - row * A.wpr;
  const int x0 = (int)(chunk * 64u + (threadIdx.x & 63u)) * CPT;
  bulk_cells<L, S, CPT>(fin, fout, kind, P, row, x0);
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
    int ix, iy, iz;
    coords(idx, P, ix, iy, iz);
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
      for (int i = 0; i < Q; i++) f[i] = S::ld_dev(fin, (int64_t)i * n + idx, L::w(i));
      float drho = 0.f, jx = 0.f, jy = 0.f, jz = 0.f;
#pragma unroll
      for (int i = 0; i < Q; i++) {
        drho += f[i];
        if (L::c(i, 0)) jx += L::c(i, 0) * f[i];
        if (L::c(i, 1)) jy += L::c(i, 1) * f[i];
        if (L::c(i, 2)) jz += L::c(i, 2) * f[i];
      }
      const float rho = 1.0f + drho;
      const float ux = jx / rho, uy = jy / rho, uz = DIM == 3 ? jz / rho : 0.0f;
      out = make_float4(ux, uy, uz, rho);
      float Fx = 0.f, Fy = 0.f, Fz = 0.f, mvSum = 0.f, mvW = 0.f, Cx = 0.f, Cy = 0.f, Cz = 0.f;
      if (kd & KIND_WALL) {
        // link sums use f + f_r - 2 w = d + d_r (deviations)
#pragma unroll
        for (int i = 1; i < Q; i++) {
          const int cx = L::c(i, 0), cy = L::c(i, 1), cz = L::c(i, 2);
          const int64_t s = safe(ix + cx, iy + cy, iz + cz);
          if (s < 0 || flags[s] == 0u) continue;
          float q = link_q(sdf, P, idx, s);
          float fr = f[i];
          if (q < 0.5f) {
            const int64_t bk = safe(ix - cx, iy - cy, iz - cz);
            if (bk >= 0 && flags[bk] == 0u) fr = 2.0f * q * f[i] + (1.0f - 2.0f * q) * S::ld_dev(fin, (int64_t)i * n + bk, L::w(i));
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
          const float m = f[i] + fr;
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
          for (int i = 0; i < Q; i++) r += S::ld_dev(fin, (int64_t)i * n + s, L::w(i));
          rs += 1.0f + r; rn += 1.0f;
        }
      }
      if (rn > 0.0f) out.w = rs / rn;
      const float4 m = meanA[idx];
      meanA[idx] = make_float4(m.x * (1.0f - a) + out.x * a, m.y * (1.0f - a) + out.y * a, m.z * (1.0f - a) + out.z * a,
                               m.w * (1.0f - a) + out.w * a);
      meanB[idx] = 0.0f;
      B3 = 1.0;
    }
    const uint32_t lo = pack_h2(out.x, out.y), hi = pack_h2(out.z, out.w);
    vel[idx] = make_ushort4((unsigned short)(lo & 0xffffu), (unsigned short)(lo >> 16), (unsigned short)(hi & 0xffffu),
                            (unsigned short)(hi >> 16));
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

// Plain streaming copy (16-byte vectors, grid-stride): the HBM-bandwidth calibration kernel.
// Each lane moves 4 independent 16-byte vectors per iteration so enough bytes are in flight.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(BLOCK) void k_copy(const u32x4* __restrict__ a, u32x4* __restrict__ b, int64_t n16) {
  const int64_t stride = (int64_t)gridDim.x * BLOCK;
  int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  for (; i + 3 * stride < n16; i += 4 * stride) {
    const u32x4 v0 = a[i], v1 = a[i + stride], v2 = a[i + 2 * stride], v3 = a[i + 3 * stride];
    __builtin_nontemporal_store(v0, &b[i]);
    __builtin_nontemporal_store(v1, &b[i + stride]);
    __builtin_nontemporal_store(v2, &b[i + 2 * stride]);
    __builtin_nontemporal_store(v3, &b[i + 3 * stride]);
  }
  for (; i < n16; i += stride) __builtin_nontemporal_store(a[i], &b[i]);
}

void launch_copy(hipStream_t st, const void* a, void* b, int64_t bytes) {
  const int64_t n16 = bytes / 16;
  const unsigned blocks = (unsigned)std::min<int64_t>((n16 + BLOCK - 1) / BLOCK, 256 * 16);
  hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(BLOCK), 0, st, (const u32x4*)a, (u32x4*)b, n16);
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

int cells_per_thread(const KParams& P) { return (P.nx % 2 == 0) ? 2 : 1; }

void launch_init(const LaunchCtx& c, void* f) {
  dispatch(c.P.q, c.prec, [&](auto l, auto s) {
    using L = decltype(l); using S = decltype(s);
    hipLaunchKernelGGL((k_init<L, S>), dim3(blocks_for(c.P.n)), dim3(BLOCK), 0, c.stream, (typename S::T*)f, c.flags, c.P);
  });
}

void launch_classify(const LaunchCtx& c, uint32_t* list, uint32_t* masks, unsigned long long* counts) {
  const int cpt = cells_per_thread(c.P);
  const int64_t ngroups = (c.P.n + cpt - 1) / cpt;
  auto go = [&](auto l) {
    using L = decltype(l);
    hipLaunchKernelGGL((k_classify<L>), dim3(blocks_for(c.P.n)), dim3(BLOCK), 0, c.stream, c.flags, c.kind, c.P);
    if (cpt == 2)
      hipLaunchKernelGGL((k_collect<L, 2>), dim3(blocks_for(ngroups)), dim3(BLOCK), 0, c.stream, c.kind, c.flags, list, masks,
                         counts, c.P);
    else
      hipLaunchKernelGGL((k_collect<L, 1>), dim3(blocks_for(ngroups)), dim3(BLOCK), 0, c.stream, c.kind, c.flags, list, masks,
                         counts, c.P);
  };
  if (c.P.q == 9) go(D2Q9{});
  else go(D3Q19{});
}

void launch_wallprep(const LaunchCtx& c, const uint32_t* list, const uint32_t* masks, WallRec* recs, int64_t count) {
  if (count <= 0) return;
  if (c.P.q == 9)
    hipLaunchKernelGGL((k_wallprep<D2Q9>), dim3(blocks_for(count)), dim3(BLOCK), 0, c.stream, c.flags, c.sdf, list, masks, recs, count, c.P);
  else
    hipLaunchKernelGGL((k_wallprep<D3Q19>), dim3(blocks_for(count)), dim3(BLOCK), 0, c.stream, c.flags, c.sdf, list, masks, recs, count, c.P);
}

StepArgs step_args(const LaunchCtx& c, const uint32_t* list, const uint32_t* masks, const WallRec* recs, int64_t count, int xcd) {
  StepArgs A;
  const int cpt = cells_per_thread(c.P);
  A.wlist = list;
  A.wmask = masks;
  A.wrec = recs;
  A.nwall = count;
  A.nwb = (uint32_t)(((count + BLOCK - 1) / BLOCK + 7) / 8 * 8);  // a multiple of 8 keeps bulk blocks XCD-aligned
  A.wpr = (uint32_t)((c.P.nx + 64 * cpt - 1) / (64 * cpt));
  A.nwaves = (uint32_t)((int64_t)c.P.ny * c.P.nz * A.wpr);
  A.nbb = (A.nwaves + 3u) / 4u;
  A.xcd = xcd;
  return A;
}

void launch_step(const LaunchCtx& c, const void* fin, void* fout, const StepArgs& A) {
  const int cpt = cells_per_thread(c.P);
  const dim3 grid(A.nwb + (A.xcd ? (A.nbb + 7u) / 8u * 8u : A.nbb));
  dispatch(c.P.q, c.prec, [&](auto l, auto s) {
    using L = decltype(l); using S = decltype(s);
    if (cpt == 2)
      hipLaunchKernelGGL((k_step<L, S, 2>), grid, dim3(BLOCK), 0, c.stream, (const typename S::T*)fin, (typename S::T*)fout, c.kind,
                         c.P, *c.W, A);
    else
      hipLaunchKernelGGL((k_step<L, S, 1>), grid, dim3(BLOCK), 0, c.stream, (const typename S::T*)fin, (typename S::T*)fout, c.kind,
                         c.P, *c.W, A);
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
