// Lattice-Boltzmann kernels for gfx950: init, cell classification, the bulk fused
// pull-stream + collide kernel, the wall kernel, and the macro / momentum-exchange reduction.
#include <hip/hip_ext.h>

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
  const uint32_t row = P.divx.div((uint32_t)idx);
  const int x = (int)((uint32_t)idx - row * (uint32_t)P.nx);
#pragma unroll
  for (int i = 0; i < L::Q; i++) {
    const float cx = (float)L::c(i, 0);
    const int64_t e = pop_at(row, i, x, L::Q, P.nx);
    if constexpr (S::BYTES == 4)
      fout[e] = L::w(i) * (1.0f + 3.0f * cx * U + 4.5f * cx * cx * U * U - 1.5f * U * U);
    else  // half-width formats store the deviation, computed directly (wgslLbm.ts:302)
      fout[e] = S::enc_dev(L::w(i) * (3.0f * cx * U + 4.5f * cx * cx * U * U - 1.5f * U * U));
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
  // the derived pass's near-wall fade of Q depends on the geometry only: solid cells along the axes
  // at one and two cells (outside the domain counts as fluid), decided here once per geometry
  auto solid_at = [&](int x, int y, int z) -> bool {
    if (x < 0 || y < 0 || z < 0 || x >= P.nx || y >= P.ny || z >= P.nz) return false;
    return flags[(int64_t)x + (int64_t)P.nx * ((int64_t)y + (int64_t)P.ny * z)] != 0u;
  };
  auto axis_hit = [&](int d) {
    bool hit = solid_at(ix + d, iy, iz) || solid_at(ix - d, iy, iz) || solid_at(ix, iy + d, iz) || solid_at(ix, iy - d, iz);
    if (L::DIM == 3) hit = hit || solid_at(ix, iy, iz + d) || solid_at(ix, iy, iz - d);
    return hit;
  };
  if (axis_hit(1)) k |= KIND_NEAR1;
  else if (axis_hit(2)) k |= KIND_NEAR2;
  kind[idx] = k;
}

// Groups of CPT cells along x (the bulk kernel's unit) that contain a WALL cell, or mix solid and
// fluid, go to the wall kernel; their cells are listed with their solid masks.
// counts[0] = listed cells, counts[2] = bulk solid cells (every other cell is a bulk fluid cell:
// the host takes their count as the remainder, which saves an atomic in nearly every wave).
// The pulled populations of a wall cell as links (see WallLink), following the reference's branch
// order for a pull (wgslLbm.ts:139-175): inlet, ground (bounce-back + belt), y-z edge mirror,
// y mirror, z periodic / mirror, solid neighbour (halfway or Bouzidi, with the fallback to q = 1/2
// when the second fluid node is missing, wgslLbm.ts:172; rotating-wheel velocity at the link
// midpoint, wgslLbm.ts:21-28), plain pull; then the per-cell wheel mass fix (wgslLbm.ts:289-291)
// folded into the constant terms. Depends on flags, distances, wheels and the parameters.
template <class L>
__device__ void wall_links(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ sdf, uint8_t kd, int64_t c,
                           int ix, int iy, int iz, uint32_t smask, const KParams& P, const KWheels& W,
                           WallLink* __restrict__ out, int64_t cap, uint32_t& meta) {
  constexpr int Q = L::Q, DIM = L::DIM;
  const int64_t n = P.n;
  const int NX = P.nx, NY = P.ny, NZ = P.nz;
  auto cell = [&](int x, int y, int z) -> int64_t { return (int64_t)x + (int64_t)NX * ((int64_t)y + (int64_t)NY * z); };
  // population element of direction d at cell (x, y, z)
  auto el = [&](int d, int x, int y, int z) -> int64_t { return pop_at((int64_t)y + (int64_t)NY * z, d, x, Q, NX); };
  const float U = P.uin;
  bool wall = false, gwall = false;
  float mvSum = 0.f, mvW = 0.f;
  float mv[Q];
  const float belt = (P.ground == 2 && !(kd & KIND_BELT_HOLD)) ? 1.0f : 0.0f;
#pragma unroll
  for (int i = 1; i < Q; i++) {
    const int cx = L::c(i, 0), cy = L::c(i, 1), cz = L::c(i, 2);
    const float w = L::w(i);
    int sx = ix - cx, sy = iy - cy, sz = DIM == 3 ? iz - cz : 0;
    if (cx < 0) sx = min(sx, NX - 1);
    const bool yOut = cy > 0 ? sy < 0 : (cy < 0 ? sy >= NY : false);
    const bool zOut = DIM == 3 ? (cz > 0 ? sz < 0 : (cz < 0 ? sz >= NZ : false)) : false;
    WallLink k;
    k.e0 = el(i, ix, iy, iz);
    k.a0 = 1.0f;
    k.a1 = 0.0f;
    k.b = 0.0f;
    k.pad = 0u;
    mv[i] = -1.0f;  // marks links without a wheel term
    if (cx > 0 && sx < 0) {
      k.a0 = 0.0f;
      k.b = inlet_dev<L>(i, U);
    } else if (cy > 0 && yOut && P.ground != 0) {
      wall = true; gwall = true;
      k.e0 = el(opp<L>(i), ix, iy, iz);
      if (cx != 0) k.b = belt * ((6.0f * w * cx) * P.uGround);
    } else {
      const int sxc = min(max(sx, 0), NX - 1);
      if (DIM == 3 && yOut && zOut) {
        k.e0 = el(mirYZ<L>(i), sxc, iy, iz);
      } else if (yOut) {
        k.e0 = el(mirY<L>(i), sxc, iy, DIM == 3 ? min(max(sz, 0), NZ - 1) : 0);
      } else if (DIM == 3 && zOut) {
        if (P.sideMode == 1) k.e0 = el(i, sxc, sy, (sz + NZ) % NZ);
        else k.e0 = el(mirZ<L>(i), sxc, sy, iz);
      } else {
        const int64_t sc = cell(sxc, sy, sz);
        if (smask & (1u << i)) {
          wall = true;
          float q = 0.5f;
          if (P.interp) {
            q = link_q(sdf, P, c, sc);
            if (q < 0.5f) {
              const int ax = ix + cx, ay = iy + cy, az = DIM == 3 ? iz + cz : 0;
              const bool ok = ax >= 0 && ay >= 0 && az >= 0 && ax < NX && ay < NY && az < NZ && flags[cell(ax, ay, az)] == 0u;
              if (!ok) q = 0.5f;
            }
          }
          k.e0 = el(opp<L>(i), ix, iy, iz);
          k.e1 = k.e0;
          if (q < 0.5f) {
            k.a0 = 2.0f * q;
            k.a1 = 1.0f - 2.0f * q;
            k.e1 = el(opp<L>(i), ix + cx, iy + cy, DIM == 3 ? iz + cz : 0);
          } else if (q > 0.5f) {
            k.a0 = 0.5f / q;
            k.a1 = 1.0f - 0.5f / q;
            k.e1 = el(i, ix, iy, iz);
          }
          const uint32_t tag = flags[sc] >> 8;
          if (tag != 0u) {
            const float3 p = make_float3(0.5f * (float)(ix + sxc) + 0.5f, 0.5f * (float)(iy + sy) + 0.5f,
                                         0.5f * (float)(iz + sz) + 0.5f);
            const float3 uw = wheel_vel(W, P, tag - 1u, p);
            const float m = (6.0f * w) * ((float)cx * uw.x + (float)cy * uw.y + (float)cz * uw.z) * (q > 0.5f ? 0.5f / q : 1.0f);
            mv[i] = m;
            mvSum += m;
            mvW += w;
          }
        } else {
          k.e0 = el(i, sxc, sy, sz);
        }
      }
    }
    if (k.a1 == 0.0f) k.e1 = k.e0;
    out[(int64_t)(i - 1) * cap] = k;
  }
  // wheel mass fix: the mean injected momentum of the wheel links is taken back from them
  const float corr = mvW > 0.0f ? mvSum / mvW : 0.0f;
#pragma unroll
  for (int i = 1; i < Q; i++)
    if (mv[i] != -1.0f) out[(int64_t)(i - 1) * cap].b = mv[i] - L::w(i) * corr;
  meta = (wall ? WM_WALL : 0u) | (gwall ? WM_GWALL : 0u) | ((uint32_t)ix << 8);
}

template <class L, int CPT>
__global__ __launch_bounds__(BLOCK) void k_collect(const uint8_t* __restrict__ kind, const uint32_t* __restrict__ flags,
                                                   uint32_t* __restrict__ list, uint32_t* __restrict__ masks,
                                                   unsigned long long* __restrict__ counts, KParams P) {
  const int64_t g = (int64_t)blockIdx.x * BLOCK + threadIdx.x;  // group index
  const int64_t ngroups = (P.n + CPT - 1) / CPT;
  int nl = 0, ns = 0;
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
    else if (alls) ns = CPT;  // else: a bulk fluid group (counted by the host as the remainder)
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
  int ts = ns;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) ts += __shfl_xor(ts, o);
  if (lane == 0) {
    if (total) base = atomicAdd(&counts[0], (unsigned long long)total);
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

// Links of the listed wall cells; rerun whenever flags, distances, wheels or parameters change.
template <class L>
__global__ __launch_bounds__(BLOCK) void k_wallprep(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ sdf,
                                                    const uint8_t* __restrict__ kind, const uint32_t* __restrict__ list,
                                                    const uint32_t* __restrict__ masks, WallLink* __restrict__ links,
                                                    uint32_t* __restrict__ meta, int64_t count, int64_t cap, KParams P, KWheels W) {
  const int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (t >= count) return;
  const int64_t c = list[t];
  int ix, iy, iz;
  coords(c, P, ix, iy, iz);
  const uint8_t kd = kind[c];
  if (kd & KIND_SOLID) {
    meta[t] = WM_SOLID;
    return;
  }
  uint32_t m = 0u;
  wall_links<L>(flags, sdf, kd, c, ix, iy, iz, masks[t], P, W, links + t, cap, m);
  meta[t] = m;
}

// ---------------------------------------------------------------------------------------------
// Bulk update. An item is a chunk of 64 * CPT cells of one x-row, owned by one wave, so the row
// (y, z) — and with it every face condition (ground, mirrors, periodic z, y-z edges) — is uniform
// across the wave and is resolved with scalar arithmetic. For each direction the wave picks the
// source direction and source row once, then every lane loads CPT consecutive cells (x shifted by
// -c_x; unaligned dword loads for FP16/BF16 pairs). Inlet (x = 0) and outlet clamp (x = nx - 1)
// are per-lane fix-ups. Groups holding a WALL cell are left to the wall path.
// An item is split in two halves so that a wave can issue the loads of its next item before it
// collides the current one (bulk_load / bulk_finish).
// HBM traffic per cell: Q loads + Q stores of S::BYTES (the algorithmic 2 Q s) + 1 kind byte.
// ---------------------------------------------------------------------------------------------
template <class L, class S, int CPT>
struct BulkIn {
  using Raw = typename std::conditional<CPT == 2, typename S::Word, typename S::T>::type;
  Raw raw[L::Q];
  uint32_t k;   // kind byte(s) of the lane's cells
  uint32_t wd;  // wall-distance byte(s) of the lane's cells (learned closure only)
};

// Row coordinates of an item (wave-uniform).
struct RowPos {
  uint32_t row;         // wave-uniform
  int iy, iz, x0;
  int xs;               // the first x of the item's chunk
  const int64_t* pull;  // the pull table (source offsets per row class)
};

// A wave-uniform base address kept opaque to the optimiser, so that the per-lane 32-bit byte offset
// is added by the memory instruction itself (global_load / global_store with a scalar base and a
// vector offset) instead of by 64-bit vector arithmetic per access, which the compiler otherwise
// emits after folding the lane offset into the sum first. Only for values the compiler holds in
// scalar registers (kernel arguments and scalar-loaded tables, not tables read after a global store).
template <class T>
__device__ __forceinline__ T* sgpr_base(T* p) {
  asm("" : "+s"(p));
  return p;
}

__device__ __forceinline__ RowPos item_pos(uint32_t item, const StepArgs& A, const KParams& P, int cpt) {
  const ItemCell c = item_cell(A, item, threadIdx.x & 63u, cpt);
  RowPos r;
  r.pull = A.pull;
  r.row = c.row;
  const uint32_t z = P.divy.div(r.row);
  r.iz = (int)z;
  r.iy = (int)(r.row - z * (uint32_t)P.ny);
  r.x0 = c.x0;
  r.xs = c.xs;
  return r;
}

// The x a lane loads for: its own, or for lanes past the row end (in the last chunk of a row whose
// length is not a multiple of the chunk) the chunk's first, so that their loads fall on the lines
// the wave reads anyway instead of on the row's first lines, which another wave reads.
__device__ __forceinline__ int load_x(const RowPos& r, int nx, int) {
  return r.x0 < nx ? r.x0 : r.xs;
}

// Row ends of a two-cells-per-lane item loaded through scalar bases (bulk_load with SADDR): the lanes
// holding a row's first or last pair read their own aligned word for the x-shifted directions and
// move the needed half into place here (wave-uniform branch: only waves holding a row end).
template <class L, class S>
__device__ __forceinline__ void bulk_load_fixup(const KParams& P, const RowPos& r, BulkIn<L, S, 2>& in) {
  using Word = typename S::Word;
  constexpr int HALF = 4 * (int)sizeof(Word);
  const int NX = P.nx;
  const int x0 = load_x(r, NX, 2);
  const bool yLo = r.iy == 0;
  const bool first = x0 == 0, last = x0 + 2 == NX;
  const int wx = __builtin_amdgcn_readfirstlane(x0);  // the wave's first x (lane 0)
  if (wx == 0 || wx + 128 >= NX) {
#pragma unroll
    for (int i = 0; i < L::Q; i++) {
      const int cx = L::c(i, 0), cy = L::c(i, 1);
      const bool grnd = cy > 0 && yLo && P.ground != 0;
      const bool edge = cx != 0 && !grnd && (cx > 0 ? first : last);
      if (cx != 0 && edge) in.raw[i] = cx > 0 ? (Word)(in.raw[i] << HALF) : (Word)(in.raw[i] >> HALF);
    }
  }
}

template <class L, class S, int CPT, bool MLP = false, bool SADDR = false, bool FIXUP = true>
__device__ __forceinline__ void bulk_load(const typename S::T* __restrict__ fin, const uint8_t* __restrict__ kind,
                                          const KParams& P, const RowPos& r, BulkIn<L, S, CPT>& in,
                                          const uint8_t* __restrict__ wdist = nullptr) {
  using T = typename S::T;
  constexpr int Q = L::Q;
  constexpr bool D3 = L::DIM == 3;
  const int NX = P.nx, NY = P.ny, NZ = P.nz;
  const int64_t n = P.n;
  const int iy = r.iy, iz = r.iz;
  // lanes past the row end load what the wave's first lane loads (in bounds, results unused)
  const int x0 = load_x(r, NX, CPT);
  const int64_t c0 = (int64_t)r.row * NX + x0;
  // every source address below is in bounds whatever the kind, so the loads need not wait for
  // the skip test
  if constexpr (CPT == 2) in.k = *glob((const uint16_t*)(kind + c0));
  else in.k = glob(kind)[c0];
  if constexpr (MLP) {
    if constexpr (CPT == 2) in.wd = *glob((const uint16_t*)(wdist + c0));
    else in.wd = glob(wdist)[c0];
  }
  const bool yLo = iy == 0, yHi = iy == NY - 1;
  const bool zLo = D3 && iz == 0, zHi = D3 && iz == NZ - 1;
  // per direction the source is the row's element plus the class's offset (wave-uniform scalars)
  const int cls = (int)yLo | ((int)yHi << 1) | ((int)zLo << 2) | ((int)zHi << 3);
  const int64_t* __restrict__ off = r.pull + cls * Q;
  const T* rowp = fin + (int64_t)r.row * Q * NX;
  const T* src[Q];
  int xlim[Q];  // largest x a lane may read (outlet clamp for CPT = 1)
#pragma unroll
  for (int i = 0; i < Q; i++) {
    src[i] = rowp + glob(off)[i];
    if constexpr (CPT == 1) {
      const int cx = L::c(i, 0), cy = L::c(i, 1);
      const bool grnd = cy > 0 && yLo && P.ground != 0;  // bounce-back pulls from the cell itself
      xlim[i] = (grnd || cx >= 0) ? NX - 1 : NX - 2;
    }
  }
  // all loads in flight together
  const uint32_t lane_off = (uint32_t)x0 * (uint32_t)sizeof(T);
  if constexpr (CPT == 2) {
    // The x-shifted word of the row's first pair (c_x > 0: cells -1, 0) and of its last pair
    // (c_x < 0: cells nx - 1, nx) reaches one element past the row, into a cache line of another
    // direction's row that nothing else of this wave reads. The element outside the row is never
    // used (the inlet replaces it, the outlet clamp reads the other half), so those two lanes read
    // the aligned word of their own pair instead and move the needed half into place (measured:
    // 4 % less read traffic at nx = 256, 5 % at nx = 512).
    using Word = typename S::Word;
    constexpr int HALF = 4 * (int)sizeof(Word);
    const int wx = __builtin_amdgcn_readfirstlane(x0);  // the wave's first x (lane 0)
    const bool first = x0 == 0, last = x0 + 2 == NX;
    if constexpr (SADDR) {
      // Scalar bases and 32-bit lane offsets (no 64-bit vector address arithmetic): the bases lie 4
      // bytes before the sources and the lane offsets 4 bytes further on, because with nx = 2 the
      // last pair starts at x0 = 0 and its aligned word lies one element before the shifted source,
      // which a non-negative offset from the shifted source could not reach. The lane offsets cost
      // a few instructions per wave; the half moves are taken only by waves holding a row end
      // (wave-uniform branch after the loads).
      const uint32_t o_mid = lane_off + 4u;
      const uint32_t o_pos = o_mid + (first ? (uint32_t)sizeof(T) : 0u);  // c_x > 0
      const uint32_t o_neg = o_mid - (last ? (uint32_t)sizeof(T) : 0u);   // c_x < 0
#pragma unroll
      for (int i = 0; i < Q; i++) {
        const int cx = L::c(i, 0), cy = L::c(i, 1);
        const bool grnd = cy > 0 && yLo && P.ground != 0;  // bounce-back: no x shift
        const uint32_t o = (cx == 0 || grnd) ? o_mid : cx > 0 ? o_pos : o_neg;
        in.raw[i] = S::ld_word((const T*)(sgpr_base((const char*)src[i] - 4) + o));
      }
      if (FIXUP) bulk_load_fixup<L, S>(P, r, in);
    } else {
      // 64-bit vector addresses; only waves holding a row end take the shifted-pair path
      // (wave-uniform branch)
      if (wx == 0 || wx + 128 >= NX) {
#pragma unroll
        for (int i = 0; i < Q; i++) {
          const int cx = L::c(i, 0), cy = L::c(i, 1);
          const bool grnd = cy > 0 && yLo && P.ground != 0;  // bounce-back: no x shift
          const bool edge = cx != 0 && !grnd && (cx > 0 ? first : last);
          // signed: with nx = 2 the last pair starts at x0 = 0 and its aligned word lies one
          // element before the shifted source (lane_off + c_x * sizeof(T) < 0)
          const int32_t o = (int32_t)lane_off + (edge ? cx * (int32_t)sizeof(T) : 0);
          const Word v = S::ld_word((const T*)((const char*)src[i] + o));
          in.raw[i] = !edge ? v : (cx > 0 ? (Word)(v << HALF) : (Word)(v >> HALF));
        }
      } else {
#pragma unroll
        for (int i = 0; i < Q; i++) in.raw[i] = S::ld_word((const T*)((const char*)src[i] + lane_off));
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < Q; i++) in.raw[i] = glob(src[i])[min(x0, xlim[i])];
  }
}

template <class L, class S, int CPT, bool REG>
__device__ __forceinline__ void bulk_finish(typename S::T* __restrict__ fout, const KParams& P, const RowPos& r,
                                            const BulkIn<L, S, CPT>& in) {
  using T = typename S::T;
  constexpr int Q = L::Q;
  const int NX = P.nx;
  const int x0 = r.x0;
  if (x0 >= NX) return;
  const int64_t n = P.n;
  const int64_t rowoff = (int64_t)r.row * Q * NX;  // the row's population block
  const uint32_t k = in.k;
  constexpr uint32_t SOLID_ALL = CPT == 2 ? 0x0101u : 0x01u;
  constexpr uint32_t SKIP = CPT == 2 ? 0x0303u : 0x03u;
  const uint32_t lane_off = (uint32_t)x0 * (uint32_t)sizeof(T);
  // (plain 64-bit address arithmetic here: with scalar bases the compiler contracted the batched
  // step's collision differently from this kernel's, and the two must agree bit for bit)
  auto store = [&](int i, float a, float b) {
    T* o = (T*)((char*)(fout + rowoff + (int64_t)i * NX) + lane_off);
    if constexpr (CPT == 2) S::st2_dev(o, L::w(i), a, b);
    else S::st_dev(o, 0, L::w(i), a);
  };
  if (k == SOLID_ALL) {
#pragma unroll
    for (int i = 0; i < Q; i++) store(i, 0.0f, 0.0f);
    return;
  }
  if (k & SKIP) return;
  const bool yLo = r.iy == 0;
  const bool gwall = yLo && P.ground != 0;
  const float U = P.uin;
  const float beltA = (P.ground == 2 && !(k & KIND_BELT_HOLD)) ? 1.0f : 0.0f;
  const float beltB = (P.ground == 2 && !((k >> 8) & KIND_BELT_HOLD)) ? 1.0f : 0.0f;
  // Pulled deviation of direction i for cell a (low halves of the loaded words) and cell b (high
  // halves), with the per-lane fix-ups: moving belt (wgslLbm.ts:159-162), inlet (wgslLbm.ts:146,
  // checked first there) and, for cell b, the outlet clamp x = nx -> nx - 1, which is the low half
  // of the same shifted word.
  auto pull_a = [&](int i) -> float {
    const int cx = L::c(i, 0), cy = L::c(i, 1);
    float v;
    if constexpr (CPT == 2) v = S::lo_dev(in.raw[i], L::w(i));
    else v = S::dev_of(in.raw[i], L::w(i));
    if (i > 0) {
      const bool grnd = cy > 0 && yLo && P.ground != 0;
      if (grnd && cx != 0) v += beltA * ((6.0f * L::w(i) * cx) * P.uGround);
      if (cx > 0 && x0 == 0) v = inlet_dev<L>(i, U);
    }
    return v;
  };
  auto pull_b = [&](int i) -> float {
    const int cx = L::c(i, 0), cy = L::c(i, 1);
    const bool grnd = cy > 0 && yLo && P.ground != 0;
    const bool clamp = i > 0 && cx < 0 && !grnd && x0 + 2 == NX;
    float v = clamp ? S::lo_dev(in.raw[i], L::w(i)) : S::hi_dev(in.raw[i], L::w(i));
    if (grnd && cx != 0) v += beltB * ((6.0f * L::w(i) * cx) * P.uGround);
    return v;
  };
  if constexpr (REG) {
    // Moments of cell a, then of cell b. The regularized outputs need only the moments, so a lane
    // carries two Relax records (not 2 Q populations) into the output loop, which writes pair by
    // pair straight to memory.
    Relax ma, mb;
    {
      float f[Q];
#pragma unroll
      for (int i = 0; i < Q; i++) f[i] = pull_a(i);
      ma = relax<L>(f, (float)x0, gwall, true, P);
      fold_inlet(ma, P);
    }
    if constexpr (CPT == 2) {
      float f[Q];
#pragma unroll
      for (int i = 0; i < Q; i++) f[i] = pull_b(i);
      mb = relax<L>(f, (float)(x0 + 1), gwall, true, P);
      fold_inlet(mb, P);
    }
    store(0, rest_out<L>(ma, true, 0.0f, P), CPT == 2 ? rest_out<L>(mb, true, 0.0f, P) : 0.0f);
#pragma unroll
    for (int i = 1; i < Q; i++) {
      const int o = opp<L>(i);
      if (o < i) continue;
      float ai, ao, bi = 0.0f, bo = 0.0f;
      pair_out<L>(ma, true, i, 0.0f, 0.0f, ai, ao, P);
      if constexpr (CPT == 2) pair_out<L>(mb, true, i, 0.0f, 0.0f, bi, bo, P);
      store(i, ai, bi);
      store(o, ao, bo);
    }
  } else {
    // BGK needs the pre-collision populations: cell a is relaxed in place, then cell b, and the
    // pairs are stored at the end.
    float fa[Q], fb[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) fa[i] = pull_a(i);
    collide_bgk<L>(fa, (float)x0, gwall, P);
    if constexpr (CPT == 2) {
#pragma unroll
      for (int i = 0; i < Q; i++) fb[i] = pull_b(i);
      collide_bgk<L>(fb, (float)(x0 + 1), gwall, P);
    }
#pragma unroll
    for (int i = 0; i < Q; i++) store(i, fa[i], CPT == 2 ? fb[i] : 0.0f);
  }
}

// Bulk finish with the learned closure (D3Q19): the moments of both cells are formed with the
// Smagorinsky tau as a first estimate, the strain rate S = -3 Pi^neq / (2 rho tau) and the other
// inputs of every cell of the wave go through the network on the MFMA path (all 64 lanes take
// part, so the skip tests come after it), and each cell relaxes with tau = t0 + 3 nu_t,
// nu_t = max(network, smag_floor x Smagorinsky); a non-finite output keeps Smagorinsky.
// Register budget (five waves per SIMD without spills): only the moments of the two cells stay
// live through the network evaluation, whose biases are read from LDS per group of 32 cells;
// the pulled values do not: the regularized outputs follow from the moments alone, and the BGK
// rows (ground rows) load their populations a second time after the network (L2 hits: the same
// wave read them just before). The collision of the row is resolved at compile time after the
// network only.
template <class L, class S, int CPT, bool REG>
__device__ __forceinline__ void bulk_finish_mlp(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                                const uint8_t* __restrict__ kind, const KParams& P, const RowPos& r,
                                                const BulkIn<L, S, CPT>& in0, const MlpDev* __restrict__ net) {
  using T = typename S::T;
  using In = BulkIn<L, S, CPT>;
  constexpr int Q = L::Q;
  const int NX = P.nx;
  const int x0 = r.x0;
  const int64_t rowoff = (int64_t)r.row * Q * NX;  // the row's population block
  const uint32_t k = in0.k;
  constexpr uint32_t SOLID_ALL = CPT == 2 ? 0x0101u : 0x01u;
  constexpr uint32_t SKIP = CPT == 2 ? 0x0303u : 0x03u;
  const uint32_t lane_off = (uint32_t)x0 * (uint32_t)sizeof(T);
  const bool yLo = r.iy == 0;
  const bool gwall = yLo && P.ground != 0;
  const bool reg_row = REG && !gwall;
  const float U = P.uin;
  const float beltA = (P.ground == 2 && !(k & KIND_BELT_HOLD)) ? 1.0f : 0.0f;
  const float beltB = (P.ground == 2 && !((k >> 8) & KIND_BELT_HOLD)) ? 1.0f : 0.0f;
  auto pull_a = [&](const In& in, int i) -> float {
    const int cx = L::c(i, 0), cy = L::c(i, 1);
    float v;
    if constexpr (CPT == 2) v = S::lo_dev(in.raw[i], L::w(i));
    else v = S::dev_of(in.raw[i], L::w(i));
    if (i > 0) {
      const bool grnd = cy > 0 && yLo && P.ground != 0;
      if (grnd && cx != 0) v += beltA * ((6.0f * L::w(i) * cx) * P.uGround);
      if (cx > 0 && x0 == 0) v = inlet_dev<L>(i, U);
    }
    return v;
  };
  auto pull_b = [&](const In& in, int i) -> float {
    const int cx = L::c(i, 0), cy = L::c(i, 1);
    const bool grnd = cy > 0 && yLo && P.ground != 0;
    const bool clamp = i > 0 && cx < 0 && !grnd && x0 + 2 == NX;
    float v = clamp ? S::lo_dev(in.raw[i], L::w(i)) : S::hi_dev(in.raw[i], L::w(i));
    if (grnd && cx != 0) v += beltB * ((6.0f * L::w(i) * cx) * P.uGround);
    return v;
  };
  auto moments_a = [&](const In& in, bool reg) {
    float f[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) f[i] = pull_a(in, i);
    return relax<L>(f, (float)x0, gwall, reg, P);
  };
  auto moments_b = [&](const In& in, bool reg) {
    float f[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) f[i] = pull_b(in, i);
    return relax<L>(f, (float)(x0 + 1), gwall, reg, P);
  };
  // inputs of both cells, normalised and packed as BF16 pairs for the wave's MFMA evaluation.
  // Cells whose result is not used (solid, wall-path or past the row end) take part unmasked: a
  // cell is one column of every product, so its inputs reach no other cell's output; a non-finite
  // input of a live cell gives a non-finite output, which keeps Smagorinsky.
  auto inputs = [&](const Relax& m, uint32_t wdb, uint32_t (&x)[MLP_IN / 2]) {
    // S = -3 Pi / (2 rho tau) = sc (-Pi), sc = 1.5 omega / rho = 6.75 omega / (4.5 rho)
    const float sc = 6.75f * m.omega * fast_rcp(m.rho45);
    float v[MLP_IN];
    mlp_step_inputs(m.pxx, m.pyy, m.pzz, m.pxy, m.pxz, m.pyz, sc, wdb, fmaf(m.ux, m.ux, fmaf(m.uy, m.uy, m.uz * m.uz)), P.inu32,
                    net->norm, v);
#pragma unroll
    for (int i = 0; i < MLP_IN / 2; i++) x[i] = pack_bf16(v[2 * i], v[2 * i + 1]);
  };
  // cell a through the network, then cell b (only one cell's inputs live at a time)
  Relax ma = moments_a(in0, reg_row), mb;
  if constexpr (CPT == 2) mb = moments_b(in0, reg_row);
  float nu_a, nu_b = 0.0f;
  {
    uint32_t x[MLP_IN / 2];
    inputs(ma, in0.wd & 0xffu, x);
    nu_a = mlp_nu_of(net, mlp_wave_z(net, x));
  }
  if constexpr (CPT == 2) {
    uint32_t x[MLP_IN / 2];
    inputs(mb, (in0.wd >> 8) & 0xffu, x);
    nu_b = mlp_nu_of(net, mlp_wave_z(net, x));
  }
  if (x0 >= NX || (k != SOLID_ALL && (k & SKIP))) return;
  // opaque copies of the source and pull-table pointers, so that the compiler neither merges the
  // BGK rows' second loads with the first ones nor keeps the first ones' values live through the
  // network
  const T* f2 = fin;
  RowPos r2 = r;
  asm volatile("" : "+s"(f2), "+s"(r2.pull));
  auto store = [&](int i, float a, float b) {
    T* o = (T*)((char*)sgpr_base(fout + rowoff + (int64_t)i * NX) + lane_off);
    if constexpr (CPT == 2) S::st2_dev(o, L::w(i), a, b);
    else S::st_dev(o, 0, L::w(i), a);
  };
  if (k == SOLID_ALL) {
#pragma unroll
    for (int i = 0; i < Q; i++) store(i, 0.0f, 0.0f);
    return;
  }
  // t0 and the inlet-layer weight again from x (cheaper than holding them through the network; the
  // opaque position keeps the compiler from reusing the first values)
  {
    float xa = (float)x0;
    asm volatile("" : "+v"(xa));
    x_terms(xa, gwall, P, ma.t0, ma.sIn);
    if constexpr (CPT == 2) x_terms(xa + 1.0f, gwall, P, mb.t0, mb.sIn);
  }
  // nu_t = max(network, floor x nu_smag), nu_smag = (tau_smag - t0) / 3; non-finite: Smagorinsky
  const float floor_ = net->norm.smag_floor;
  auto finish = [&](auto rg) {
    constexpr bool reg = decltype(rg)::value;
    In in;
    if constexpr (!reg) bulk_load<L, S, CPT, false, false>(f2, kind, P, r2, in);
    else in = in0;  // unused: the regularized outputs follow from the moments
    // t0 + 3 max(nu, floor nu_smag) = max(t0 + 3 nu, t0 + floor (tau_smag - t0)); the comparison is
    // false for NaN as well
    auto closure_tau = [&](Relax& m, float nu) {
      if (nu < 1e3f) retau(m, fmaxf(fmaf(3.0f, nu, m.t0), fmaf(floor_, fast_rcp(m.omega) - m.t0, m.t0)), reg);
    };
    closure_tau(ma, nu_a);
    if constexpr (CPT == 2) closure_tau(mb, nu_b);
    if constexpr (reg) {
      fold_inlet(ma, P);
      if constexpr (CPT == 2) fold_inlet(mb, P);
    }
    store(0, rest_out<L>(ma, reg, pull_a(in, 0), P), CPT == 2 ? rest_out<L>(mb, reg, pull_b(in, 0), P) : 0.0f);
#pragma unroll
    for (int i = 1; i < Q; i++) {
      const int o = opp<L>(i);
      if (o < i) continue;
      float ai, ao, bi = 0.0f, bo = 0.0f;
      pair_out<L>(ma, reg, i, pull_a(in, i), pull_a(in, o), ai, ao, P);
      if constexpr (CPT == 2) pair_out<L>(mb, reg, i, pull_b(in, i), pull_b(in, o), bi, bo, P);
      store(i, ai, bi);
      store(o, ao, bo);
    }
  };
  if (reg_row) finish(std::true_type{});
  else finish(std::false_type{});
}

// One 32-byte wall link, read as two 16-byte global loads.
__device__ __forceinline__ WallLink load_link(const WallLink* p) {
  typedef uint32_t u32x4v __attribute__((ext_vector_type(4)));
  struct alignas(16) Pair { u32x4v a, b; } v;
  const u32x4v* q = (const u32x4v*)p;
  v.a = *glob(q);
  v.b = *glob(q + 1);
  return __builtin_bit_cast(WallLink, v);
}

// Wall cells: one lane per listed cell (cells next to the body). Every pulled population is a
// precomputed link (two independent gathers and two FMAs), so all loads of a lane are in flight
// together; then the collision (BGK on wall cells, wgslLbm.ts:186).
template <class L, class S>
__device__ __forceinline__ void wall_cell(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                          const StepArgs& A, const KParams& P, int64_t t) {
  constexpr int Q = L::Q;
  if (t >= A.nwall) return;
  const int64_t c = glob(A.wlist)[t];
  const uint32_t meta = glob(A.wmeta)[t];
  const uint32_t row = P.divx.div((uint32_t)c);
  const int64_t base = pop_at(row, 0, (int)((uint32_t)c - row * (uint32_t)P.nx), Q, P.nx);  // direction i: base + i * nx
  const int64_t nx = P.nx;
  float f[Q];
  if (meta & WM_SOLID) {
#pragma unroll
    for (int i = 0; i < Q; i++) S::st_dev(fout, base + i * nx, L::w(i), 0.0f);
    return;
  }
  f[0] = S::ld_dev(fin, base, L::w(0));
#pragma unroll
  for (int i = 1; i < Q; i++) {
    const WallLink k = load_link(A.wlink + (int64_t)(i - 1) * A.wcap + t);
    const float x0 = S::ld_dev(fin, k.e0, L::w(i));
    const float x1 = S::ld_dev(fin, k.e1, L::w(i));
    f[i] = fmaf(k.a0, x0, fmaf(k.a1, x1, k.b));
  }
  collide<L>(f, (float)(meta >> 8), (meta & WM_WALL) != 0u, (meta & WM_GWALL) != 0u, P);
#pragma unroll
  for (int i = 0; i < Q; i++) S::st_dev(fout, base + i * nx, L::w(i), f[i]);
}

// One lattice step in one launch. A.nwb blocks update the wall cells (one lane per cell, gathers
// from the precomputed links), placed among the bulk blocks beside the rows they read (block_role);
// every other block is 4 waves of the bulk update, one item per wave. Items are numbered (row, chunk) with the chunk
// fastest. Bulk block b runs on XCD b % 8 (blocks are dealt round-robin to the 8 XCDs); the items
// are cut into runs of A.xk consecutive items, the runs dealt round-robin to the XCDs, so that
// the cache lines shared by neighbouring rows of one run stay inside one L2. A.xk = 4 is the plain
// order (block b takes items 4b..4b+3); A.xk = nwaves / 8 gives every XCD one contiguous eighth.
// Occupancy of the plain step: five waves per SIMD for the 16-bit formats measured faster than six
// (419.7 vs 426.5 us at 512x256x256, 57.9 vs 59.7 us at 256x128x128) and than four (425) or seven
// (436, with spills).
template <class S> constexpr int step_waves() { return S::BYTES == 2 ? 5 : 4; }
// The batched step reads its parameters and buffers from the descriptor table, which costs it
// registers: at six waves per SIMD its 16-bit forms spilled (32 bytes per lane); five waves keep
// them spill-free and measured faster on the 24-domain sweep (61.2 vs 60.4 k MLUPS).
template <class S> constexpr int batch_waves() { return S::BYTES == 2 ? 5 : 4; }

// The bulk item of wave w of launch block b (see k_step for the placement).
__device__ __forceinline__ uint32_t bulk_item(const StepArgs& A, uint32_t b, uint32_t w) {
  const uint32_t lw = (b / 8u) * 4u + w;  // wave index on this XCD
  const uint32_t m = A.dxk.div(lw);
  return (m * 8u + b % 8u) * A.xk + (lw - m * A.xk);
}

template <class L, class S, int CPT, bool REG>
__device__ __forceinline__ void bulk_block(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                           const uint8_t* __restrict__ kind, const KParams& P, const StepArgs& A, uint32_t b) {
  const uint32_t item = bulk_item(A, b, __builtin_amdgcn_readfirstlane(threadIdx.x >> 6));
  if (item >= A.nwaves) return;
  BulkIn<L, S, CPT> in;
  const RowPos r = item_pos(item, A, P, CPT);
  bulk_load<L, S, CPT>(fin, kind, P, r, in);
  // ground rows relax with BGK (wall cells, wgslLbm.ts:186); the branch is wave-uniform
  if (REG && !(r.iy == 0 && P.ground != 0)) bulk_finish<L, S, CPT, true>(fout, P, r, in);
  else bulk_finish<L, S, CPT, false>(fout, P, r, in);
}

// The learned step's bulk block. The network is staged in LDS by the whole block; each wave reads
// its piece of the network first and issues its population loads before the piece is written and
// the block synchronises, so the staging's round trip overlaps the populations' instead of preceding
// it (the piece's load is issued first, so the LDS write's wait leaves the population loads in flight).
// The population loads use scalar bases (fewer vector instructions where the vector pipe is the
// bound); the plain step keeps 64-bit vector addresses, measured faster at 512x256x256.
template <class S, int CPT, bool REG>
__device__ __forceinline__ void bulk_block_mlp(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                               const uint8_t* __restrict__ kind, const KParams& P, const StepArgs& A, uint32_t b,
                                               const uint8_t* __restrict__ wdist, const MlpDev* __restrict__ net, MlpDev* snet) {
  using L = D3Q19;
  const uint32_t item = bulk_item(A, b, __builtin_amdgcn_readfirstlane(threadIdx.x >> 6));
  const bool live = item < A.nwaves;  // wave-uniform; every wave takes part in the staging
  const u32x4_hw piece = stage_net_load(net);
  // waves past the last item load the last item too (in bounds, results unused): with the loads in
  // a branch the LDS write after it would wait for every load of the taken side
  BulkIn<L, S, CPT> in;
  const RowPos r = item_pos(live ? item : A.nwaves - 1u, A, P, CPT);
  bulk_load<L, S, CPT, true, true, false>(fin, kind, P, r, in, wdist);
  stage_net_store(snet, piece);
  if (!live) return;
  if constexpr (CPT == 2) bulk_load_fixup<L, S>(P, r, in);  // after the barrier: it waits for the loads
  bulk_finish_mlp<L, S, CPT, REG>(fin, fout, kind, P, r, in, snet);
}

// What block b of a step launch does: launch blocks come in groups of 8 (one per XCD), each group
// 8 wall blocks or 8 bulk blocks in the order of the role table, which puts every wall group just
// after the bulk group that streams the rows its links gather from (see build_roles): the wall
// cells' scattered reads and partial-line writes then meet lines the bulk blocks moved through the
// caches moments before instead of going to HBM on their own. The index keeps b's XCD (b % 8).
struct BlockRole {
  bool wall;
  uint32_t index;  // wall block or bulk block index
};
__device__ __forceinline__ BlockRole block_role(const StepArgs& A, uint32_t b) {
  const uint32_t r = __builtin_amdgcn_readfirstlane(glob(A.role)[b >> 3]);
  return {(r >> 31) != 0u, (r & 0x7fffffffu) * 8u + (b & 7u)};
}

template <class L, class S, int CPT, bool REG>
__global__ __launch_bounds__(BLOCK, step_waves<S>()) void k_step(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                                const uint8_t* __restrict__ kind, KParams P, StepArgs A) {
  const BlockRole br = block_role(A, blockIdx.x);
  if (br.wall) {
    wall_cell<L, S>(fin, fout, A, P, (int64_t)br.index * BLOCK + threadIdx.x);
    return;
  }
  bulk_block<L, S, CPT, REG>(fin, fout, kind, P, A, br.index);
}

// The step with the learned closure in the bulk (D3Q19); wall cells keep Smagorinsky.
template <class S, int CPT, bool REG>
__global__ __launch_bounds__(BLOCK, step_waves<S>()) void k_step_mlp(const typename S::T* __restrict__ fin, typename S::T* __restrict__ fout,
                                                    const uint8_t* __restrict__ kind, const uint8_t* __restrict__ wdist,
                                                    const MlpDev* __restrict__ net, KParams P, StepArgs A) {
  __shared__ MlpDev snet;
  const BlockRole br = block_role(A, blockIdx.x);
  if (br.wall) {
    wall_cell<D3Q19, S>(fin, fout, A, P, (int64_t)br.index * BLOCK + threadIdx.x);
    return;
  }
  bulk_block_mlp<S, CPT, REG>(fin, fout, kind, P, A, br.index, wdist, net, &snet);  // bulk blocks only: the whole block
}

// One lattice step of many independent domains in one launch (sweeps): the same wall and bulk
// work as k_step, each block finding its domain through `dmap` (block -> domain) and its
// parameters, layout, buffers and role table in the descriptor table. Each domain's blocks form
// one contiguous run starting at a multiple of 8.
template <class L, class S, int CPT, bool REG>
__global__ __launch_bounds__(BLOCK, batch_waves<S>()) void k_step_batch(const DomainDesc* __restrict__ D, const uint32_t* __restrict__ dmap,
                                                                      int32_t parity) {
  using T = typename S::T;
  const uint32_t b = blockIdx.x;
  const uint32_t d = __builtin_amdgcn_readfirstlane(dmap[b]);
  const DomainDesc& dd = D[d];
  // (the pointers read from the table are accessed through glob(), see lbm.hpp)
  const T* fin = (const T*)dd.f[parity];
  T* fout = (T*)dd.f[parity ^ 1];
  const BlockRole br = block_role(dd.A, b - dd.first);
  if (br.wall) {
    wall_cell<L, S>(fin, fout, dd.A, dd.P, (int64_t)br.index * BLOCK + threadIdx.x);
    return;
  }
  bulk_block<L, S, CPT, REG>(fin, fout, dd.kind, dd.P, dd.A, br.index);
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
    auto el = [&](int d, int x, int y, int z) -> int64_t { return pop_at((int64_t)y + (int64_t)NY * z, d, x, Q, NX); };
    float4 out = make_float4(0.f, 0.f, 0.f, 1.f);
    const uint8_t kd = kind[idx];
    const float a = P.emaAlpha;
    if (!(kd & KIND_SOLID)) {
      float f[Q];
#pragma unroll
      for (int i = 0; i < Q; i++) f[i] = S::ld_dev(fin, el(i, ix, iy, iz), L::w(i));
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
            if (bk >= 0 && flags[bk] == 0u)
              fr = 2.0f * q * f[i] + (1.0f - 2.0f * q) * S::ld_dev(fin, el(i, ix - cx, iy - cy, iz - cz), L::w(i));
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
          for (int i = 0; i < Q; i++) r += S::ld_dev(fin, el(i, ix + nb[k][0], iy + nb[k][1], iz + nb[k][2]), L::w(i));
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
  // block reduction: waves via cross-lane shuffles, then LDS across the 4 waves. The counts
  // (reference-column cells, bad cells, solid cells) are popcounts of the wave's ballot, exact; the
  // sums (forces on the body's links, density over the reference column) are reduced only in waves
  // where some lane holds a non-zero term: a wave of zero terms sums to zero either way, and the
  // other waves add in the same order as before
  B1 = (double)__popcll(__ballot(B1 != 0.0));
  B2 = (double)__popcll(__ballot(B2 != 0.0));
  B3 = (double)__popcll(__ballot(B3 != 0.0));
  if (__any(F0 != 0.0 || F1 != 0.0 || F2 != 0.0)) {
    F0 = wave_sum(F0); F1 = wave_sum(F1); F2 = wave_sum(F2);
  } else {
    F0 = F1 = F2 = 0.0;
  }
  B0 = __any(B0 != 0.0) ? wave_sum(B0) : 0.0;
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

// Fixed-order reduction of `count` records of 8 doubles (sums; slot 3 a maximum) into out[8] by one
// block of T threads: thread t folds records t, t + T, ... in order, then each wave (cross-lane, a
// fixed tree) and the block (LDS, in wave order). The result depends only on count.
template <int T>
__device__ __forceinline__ void reduce_records(const double* __restrict__ p, int64_t count, double* __restrict__ out) {
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t b = threadIdx.x; b < count; b += T) {
#pragma unroll
    for (int k = 0; k < 8; k++) acc[k] = (k == 3) ? fmax(acc[k], p[b * 8 + k]) : acc[k] + p[b * 8 + k];
  }
  __shared__ double sh[T / 64][8];
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
    for (int w = 1; w < T / 64; w++) v = (k == 3) ? fmax(v, sh[w][k]) : v + sh[w][k];
    out[k] = v;
  }
}

// The block partials are reduced in two stages (wgslLbm.ts:491-521 reduces them in one workgroup):
// chunks of REDUCE_CHUNK records by one 256-thread block each, then the chunk results by one block.
// A single block reading all of them (8 MB at 512x256x256) is bound by one CU's bandwidth (196 us).
constexpr int64_t REDUCE_CHUNK = 512;
__host__ __device__ constexpr int64_t reduce_chunks(int64_t records) { return (records + REDUCE_CHUNK - 1) / REDUCE_CHUNK; }

__global__ __launch_bounds__(256) void k_reduce_chunks(const double* __restrict__ partials, int64_t count, double* __restrict__ stage) {
  const int64_t first = (int64_t)blockIdx.x * REDUCE_CHUNK;
  const int64_t m = count - first < REDUCE_CHUNK ? count - first : REDUCE_CHUNK;
  reduce_records<256>(partials + first * 8, m, stage + (int64_t)blockIdx.x * 8);
}

// Final reduction into result[8], FP64.
__global__ __launch_bounds__(1024) void k_reduce(const double* __restrict__ partials, int64_t count, double* __restrict__ result) {
  reduce_records<1024>(partials, count, result);
}

// Convert between device storage and full-value host floats. The host side is direction-major,
// out[i * n + cell]; element e of it maps to the row-blocked device element pop_at(row, i, x).
__device__ inline int64_t host_to_dev(int64_t e, const KParams& P, int q, int& i) {
  i = (int)(e / P.n);
  const int64_t c = e - (int64_t)i * P.n;
  const uint32_t row = P.divx.div((uint32_t)c);
  return pop_at(row, i, (int)((uint32_t)c - row * (uint32_t)P.nx), q, P.nx);
}
template <class L, class S>
__global__ __launch_bounds__(BLOCK) void k_unpack(const typename S::T* __restrict__ f, float* __restrict__ out, KParams P) {
  const int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (e >= (int64_t)L::Q * P.n) return;
  int i;
  const int64_t d = host_to_dev(e, P, L::Q, i);
  float w = 0.0f;
#pragma unroll
  for (int k = 0; k < L::Q; k++) if (k == i) w = L::w(k);
  out[e] = S::ld(f, d, w);
}
// The cells of planes z0 .. z0 + count - 1 as full values, direction-major within the slab:
// out[i * m + c], m = nx * ny * count, c = x + nx * (y + ny * (z - z0)).
template <class L, class S>
__global__ __launch_bounds__(BLOCK) void k_unpack_planes(const typename S::T* __restrict__ f, float* __restrict__ out, KParams P,
                                                         int32_t z0, int64_t m) {
  const int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (e >= (int64_t)L::Q * m) return;
  const int i = (int)(e / m);
  const int64_t cell = e - (int64_t)i * m + (int64_t)z0 * P.nx * P.ny;
  const int64_t row = cell / P.nx;
  const int x = (int)(cell - row * P.nx);
  float w = 0.0f;
#pragma unroll
  for (int k = 0; k < L::Q; k++) if (k == i) w = L::w(k);
  out[e] = S::ld(f, pop_at(row, i, x, L::Q, P.nx), w);
}

template <class L, class S>
__global__ __launch_bounds__(BLOCK) void k_pack(const float* __restrict__ in, typename S::T* __restrict__ f, KParams P) {
  const int64_t e = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (e >= (int64_t)L::Q * P.n) return;
  int i;
  const int64_t d = host_to_dev(e, P, L::Q, i);
  float w = 0.0f;
#pragma unroll
  for (int k = 0; k < L::Q; k++) if (k == i) w = L::w(k);
  f[d] = S::enc(in[e], w);
}

// Plain streaming copy, one 16-byte vector per lane over a grid that covers the buffer: the
// HBM-bandwidth calibration kernel (the fastest of the simple copy shapes on MI355X).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(BLOCK) void k_copy(const u32x4* __restrict__ a, u32x4* __restrict__ b, int64_t n16) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n16) __builtin_nontemporal_store(a[i], &b[i]);
}

void launch_copy(hipStream_t st, const void* a, void* b, int64_t bytes) {
  const int64_t n16 = bytes / 16;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)((n16 + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, st, (const u32x4*)a, (u32x4*)b, n16);
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

int cells_per_thread(const KParams& P) { return (P.cpt == 2 && P.nx % 2 == 0) ? 2 : 1; }

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

void launch_wallprep(const LaunchCtx& c, const uint32_t* list, const uint32_t* masks, WallLink* links, uint32_t* meta, int64_t count,
                     int64_t cap) {
  if (count <= 0) return;
  if (c.P.q == 9)
    hipLaunchKernelGGL((k_wallprep<D2Q9>), dim3(blocks_for(count)), dim3(BLOCK), 0, c.stream, c.flags, c.sdf, c.kind, list, masks, links,
                       meta, count, cap, c.P, *c.W);
  else
    hipLaunchKernelGGL((k_wallprep<D3Q19>), dim3(blocks_for(count)), dim3(BLOCK), 0, c.stream, c.flags, c.sdf, c.kind, list, masks, links,
                       meta, count, cap, c.P, *c.W);
}

int pull_table_len(int q) { return PULL_CLASSES * q; }

void pull_table(const KParams& P, int64_t* out) {
  auto go = [&](auto l) {
    using L = decltype(l);
    const int64_t nx = P.nx;
    for (int cls = 0; cls < PULL_CLASSES; cls++)
      for (int i = 0; i < L::Q; i++) {
        int d, dy, dz, dx;
        pull_source<L>(i, (cls & 1) != 0, (cls & 2) != 0, (cls & 4) != 0, (cls & 8) != 0, P.ground, P.sideMode, P.nz, d, dy, dz, dx);
        // the source cell's row is row + dy + dz * ny; its block starts (dy + dz * ny) * Q * nx later
        out[cls * L::Q + i] = (((int64_t)dy + (int64_t)dz * P.ny) * L::Q + d) * nx + dx;
      }
  };
  if (P.q == 9) go(D2Q9{});
  else go(D3Q19{});
}

static_assert(WALL_GROUP_CELLS == 8 * BLOCK, "a wall group is 8 wall blocks");

std::vector<uint32_t> build_roles(const StepArgs& A, const std::vector<uint32_t>& wall_rows, int ny, int nz) {
  const uint32_t nbg = A.nbb / 8, nwg = A.nwb / 8;
  const int64_t rows = (int64_t)ny * nz;
  // bulk group (launch order, per XCD) that holds the last item of each wall group's target row
  std::vector<std::vector<uint32_t>> after(nbg + 1);
  for (uint32_t k = 0; k < nwg; k++) {
    const int64_t row = k < wall_rows.size() ? std::min<int64_t>((int64_t)wall_rows[k] + ny + 1, rows - 1) : rows - 1;
    const int64_t item = std::min<int64_t>(last_item_of_row(A, row), (int64_t)A.nwaves - 1);
    const int64_t run = item / A.xk;  // runs of xk items, dealt round-robin to the XCDs (bulk_block)
    const int64_t lw = (run / 8) * A.xk + item % A.xk;  // wave index on its XCD
    after[std::min<int64_t>(lw / 4, nbg)].push_back(k);  // bulk group = (bulk block) / 8 = lw / 4
  }
  std::vector<uint32_t> role;
  role.reserve(nbg + nwg);
  for (uint32_t g = 0; g <= nbg; g++) {
    if (g < nbg) role.push_back(g);
    for (uint32_t k : after[g]) role.push_back(0x80000000u | k);
  }
  return role;
}

StepArgs step_args(const LaunchCtx& c, const uint32_t* list, const uint32_t* meta, const WallLink* links, int64_t count, int64_t cap,
                   int64_t chunk) {
  StepArgs A{};
  const int cpt = cells_per_thread(c.P);
  A.pull = c.pull;
  A.wlist = list;
  A.wmeta = meta;
  A.wlink = links;
  A.wcap = cap;
  A.nwall = count;
  A.nwb = (uint32_t)(((count + BLOCK - 1) / BLOCK + 7) / 8 * 8);  // a multiple of 8 keeps bulk blocks XCD-aligned
  set_item_layout(A, c.P.nx, c.P.ny, c.P.nz, cpt);
  A.pad = 0;
  // chunk <= 0: one contiguous eighth of the items per XCD
  const int64_t per = ((int64_t)A.nwaves + 7) / 8;
  A.xk = (uint32_t)std::max<int64_t>(1, std::min<int64_t>(chunk > 0 ? chunk : per, per));
  A.dxk.init(A.xk);
  const int64_t runs = ((int64_t)A.nwaves + A.xk - 1) / A.xk;
  const int64_t waves_per_xcd = (runs + 7) / 8 * (int64_t)A.xk;
  A.nbb = (uint32_t)(8 * ((waves_per_xcd + 3) / 4));
  return A;
}

void launch_step_batch(hipStream_t st, int q, int prec, int cpt, bool reg, const DomainDesc* descs, const uint32_t* dmap,
                       uint32_t blocks, int parity) {
  dispatch(q, prec, [&](auto l, auto s) {
    using L = decltype(l); using S = decltype(s);
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(blocks), dim3(BLOCK), 0, st, descs, dmap, parity); };
    // 16-bit words hold two cells: those formats always run two cells per lane
    if (S::BYTES == 2 || cpt == 2) {
      if (reg) go(k_step_batch<L, S, 2, true>);
      else go(k_step_batch<L, S, 2, false>);
    } else if constexpr (S::BYTES == 4) {
      if (reg) go(k_step_batch<L, S, 1, true>);
      else go(k_step_batch<L, S, 1, false>);
    }
  });
}

void launch_step(const LaunchCtx& c, const void* fin, void* fout, const StepArgs& A, hipEvent_t ev0, hipEvent_t ev1) {
  const int cpt = cells_per_thread(c.P);
  const bool reg = c.P.collision != 1;
  const dim3 grid(A.nwb + A.nbb);
  if (c.P.closure == 1 && c.mlp && c.wdist && c.P.q == 19) {
    auto go = [&](auto kern, auto s) {
      using S = decltype(s);
      hipExtLaunchKernelGGL(kern, grid, dim3(BLOCK), 0, c.stream, ev0, ev1, 0, (const typename S::T*)fin, (typename S::T*)fout, c.kind,
                            c.wdist, c.mlp, c.P, A);
    };
    auto pick = [&](auto s) {
      using S = decltype(s);
      if (S::BYTES == 2 || cpt == 2) {
        if (reg) go(k_step_mlp<S, 2, true>, s);
        else go(k_step_mlp<S, 2, false>, s);
      } else if constexpr (S::BYTES == 4) {
        if (reg) go(k_step_mlp<S, 1, true>, s);
        else go(k_step_mlp<S, 1, false>, s);
      }
    };
    if (c.prec == 0) pick(SF32{});
    else if (c.prec == 1) pick(SF16{});
    else pick(SBF16{});
    return;
  }
  dispatch(c.P.q, c.prec, [&](auto l, auto s) {
    using L = decltype(l); using S = decltype(s);
    // with events: hipExtLaunchKernelGGL records them at the kernel's own start and end
    auto go = [&](auto kern) {
      hipExtLaunchKernelGGL(kern, grid, dim3(BLOCK), 0, c.stream, ev0, ev1, 0, (const typename S::T*)fin, (typename S::T*)fout, c.kind,
                            c.P, A);
    };
    if (S::BYTES == 2 || cpt == 2) {
      if (reg) go(k_step<L, S, 2, true>);
      else go(k_step<L, S, 2, false>);
    } else if constexpr (S::BYTES == 4) {
      if (reg) go(k_step<L, S, 1, true>);
      else go(k_step<L, S, 1, false>);
    }
  });
}

int64_t macro_records(int64_t n) {
  const int64_t blocks = (int64_t)blocks_for(n);
  return blocks + reduce_chunks(blocks);
}

void launch_macro(const LaunchCtx& c, const void* fin, void* vel, void* meanA, float* meanB, double* partials, double* result) {
  dispatch(c.P.q, c.prec, [&](auto l, auto s) {
    using L = decltype(l); using S = decltype(s);
    hipLaunchKernelGGL((k_macro<L, S>), dim3(blocks_for(c.P.n)), dim3(BLOCK), 0, c.stream, (const typename S::T*)fin, c.flags,
                       c.sdf, c.kind, (ushort4*)vel, (float4*)meanA, meanB, partials, c.P, *c.W);
  });
  // the block partials, then the chunk results right after them in the same buffer (macro_records)
  const int64_t blocks = (int64_t)blocks_for(c.P.n), chunks = reduce_chunks(blocks);
  double* stage = partials + blocks * 8;
  hipLaunchKernelGGL(k_reduce_chunks, dim3((uint32_t)chunks), dim3(256), 0, c.stream, (const double*)partials, blocks, stage);
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, c.stream, (const double*)stage, chunks, result);
}

void launch_unpack(const LaunchCtx& c, const void* f, float* out) {
  dispatch(c.P.q, c.prec, [&](auto l, auto s) {
    using L = decltype(l); using S = decltype(s);
    hipLaunchKernelGGL((k_unpack<L, S>), dim3(blocks_for((int64_t)L::Q * c.P.n)), dim3(BLOCK), 0, c.stream,
                       (const typename S::T*)f, out, c.P);
  });
}

void launch_unpack_planes(const LaunchCtx& c, const void* f, int32_t z0, int32_t count, float* out) {
  const int64_t m = (int64_t)c.P.nx * c.P.ny * count;
  dispatch(c.P.q, c.prec, [&](auto l, auto s) {
    using L = decltype(l); using S = decltype(s);
    hipLaunchKernelGGL((k_unpack_planes<L, S>), dim3(blocks_for((int64_t)L::Q * m)), dim3(BLOCK), 0, c.stream,
                       (const typename S::T*)f, out, c.P, z0, m);
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
