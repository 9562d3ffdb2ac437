// Device-side building blocks of the lattice-Boltzmann hot path for gfx950 (MI355X).
//
// Layout in HBM (all buffers owned by the solver handle):
//   populations  row-blocked: f[(row * Q + i) * nx + x], row = y + ny * z (see pop_at).
//                FP32 stores full f_i; FP16 / BF16 store the deviation f_i - w_i (RNE), so two
//                x-neighbouring cells form one 32-bit word exactly like the reference's
//                pack2x16float pairs (src/solver/wgslLbm.ts:62-75).
//   flags        u32 per cell: 0 fluid, 1 solid, 1 | (k+1) << 8 solid of wheel k (types as in
//                src/solver/SolverGPU.ts:341-344).
//   sdf          u32 per cell: ~bits(distance) or 0 = unknown (SolverGPU.ts:312-322).
//   kind         u8 per cell, derived from flags (see KIND_* below).
//
// Arithmetic is FP32 on deviations d_i = f_i - w_i: moments are formed from d (rho = 1 + sum d,
// j = sum c d, second moment = sum c c d + delta/3), which keeps the round-off relative to the
// (small) deviations instead of to f.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wt {

// Device pointers handed to the kernels always address global memory. Accesses go through
// glob(), a cast to the global address space: pointers the compiler cannot trace to a kernel
// argument (the batched step reads them from its domain table) would otherwise be accessed with
// flat instructions, which also count against the scalar-load wait counter and serialise every
// parameter fetch behind the population loads.
#define WT_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const WT_GLOBAL T* glob(const T* p) { return (const WT_GLOBAL T*)p; }
template <class T>
__device__ __forceinline__ WT_GLOBAL T* glob(T* p) { return (WT_GLOBAL T*)p; }

// Population layout, row-blocked: the Q values of one x-row of cells (row = y + ny * z) lie in Q
// consecutive runs of nx elements, f[(row * Q + i) * nx + x]. A wave's pulls for one row then touch
// a few neighbouring row blocks instead of Q planes n elements apart (fewer open DRAM pages; 7 %
// faster than direction-major planes on MI355X at 256x128x128 and 512x256x256).
__host__ __device__ inline int64_t pop_at(int64_t row, int i, int x, int q, int nx) { return (row * q + i) * (int64_t)nx + x; }

enum : uint8_t {
  KIND_FLUID = 0,      // fluid cell whose plain pulls all come from fluid cells: bulk kernel
  KIND_SOLID = 1,      // solid cell (rest state)
  KIND_WALL = 2,       // fluid cell with at least one plain pull from a solid cell: wall kernel
  KIND_BELT_HOLD = 4,  // floor cell where a tyre contact patch holds the belt still (wgslLbm.ts:276-287)
  KIND_NEAR1 = 8,      // fluid cell with a solid cell one cell away along an axis (the Q-criterion fade)
  KIND_NEAR2 = 16,     // fluid cell with a solid cell two cells away along an axis and none at one
};

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
// Population storage formats. `*_dev` accessors move deviations d = f - w.
// ---------------------------------------------------------------------------------------------
typedef uint32_t u32_a2 __attribute__((aligned(2)));
typedef uint64_t u64_a4 __attribute__((aligned(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x2a4 __attribute__((ext_vector_type(2), aligned(4)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// packed conversions (v_cvt_pk_f16_f32 / v_cvt_pk_bf16_f32 on gfx950; round to nearest even)
__device__ inline uint32_t pack_h2(float a, float b) { f32x2 v = {a, b}; return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, f16x2)); }
__device__ inline uint32_t pack_b2(float a, float b) { f32x2 v = {a, b}; return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2)); }
__device__ inline void unpack_h2(uint32_t u, float& a, float& b) {
  f32x2 v = __builtin_convertvector(__builtin_bit_cast(f16x2, u), f32x2);
  a = v.x; b = v.y;
}
__device__ inline void unpack_b2(uint32_t u, float& a, float& b) {
  a = __uint_as_float(u << 16);
  b = __uint_as_float(u & 0xffff0000u);
}
__device__ inline float h2f(uint32_t h) { float a, b; unpack_h2(h, a, b); return a; }
__device__ inline float b2f(uint32_t h) { return __uint_as_float(h << 16); }

struct SF32 {
  using T = float;
  static constexpr int BYTES = 4;
  __device__ static float ld(const T* p, int64_t e, float) { return glob(p)[e]; }
  __device__ static float ld_dev(const T* p, int64_t e, float w) { return glob(p)[e] - w; }
  __device__ static T enc(float f, float) { return f; }
  __device__ static void st_dev(T* p, int64_t e, float w, float d) { glob(p)[e] = w + d; }
  // two consecutive cells; FP32 pairs are only 4-byte aligned in general
  template <bool ALIGNED>
  __device__ static void ld2_dev(const T* p, float w, float& a, float& b) {
    const uint64_t v = *(const u64_a4*)p;
    a = __uint_as_float((uint32_t)v) - w;
    b = __uint_as_float((uint32_t)(v >> 32)) - w;
  }
  __device__ static void st2_dev(T* p, float w, float a, float b) {
    f32x2a4 v = {w + a, w + b};
    __builtin_nontemporal_store(v, glob((f32x2a4*)p));
  }
  // raw access for the bulk kernel: a word holds two consecutive cells
  using Word = uint64_t;
  __device__ static Word ld_word(const T* p) { return *glob((const u64_a4*)p); }
  __device__ static void unpack_word(Word v, float w, float& a, float& b) {
    a = __uint_as_float((uint32_t)v) - w;
    b = __uint_as_float((uint32_t)(v >> 32)) - w;
  }
  __device__ static float lo_dev(Word v, float w) { return __uint_as_float((uint32_t)v) - w; }
  __device__ static float hi_dev(Word v, float w) { return __uint_as_float((uint32_t)(v >> 32)) - w; }
  __device__ static float dev_of(T v, float w) { return v - w; }
};

struct SF16 {
  using T = uint16_t;
  static constexpr int BYTES = 2;
  __device__ static float ld(const T* p, int64_t e, float w) { return h2f(glob(p)[e]) + w; }
  __device__ static float ld_dev(const T* p, int64_t e, float) { return h2f(glob(p)[e]); }
  __device__ static T enc(float f, float w) { return (T)(pack_h2(f - w, 0.0f) & 0xffffu); }
  __device__ static T enc_dev(float d) { return (T)(pack_h2(d, 0.0f) & 0xffffu); }
  __device__ static void st_dev(T* p, int64_t e, float, float d) { glob(p)[e] = enc_dev(d); }
  template <bool ALIGNED>
  __device__ static void ld2_dev(const T* p, float, float& a, float& b) {
    const uint32_t v = ALIGNED ? *(const uint32_t*)p : *(const u32_a2*)p;
    unpack_h2(v, a, b);
  }
  __device__ static void st2_dev(T* p, float, float a, float b) { __builtin_nontemporal_store(pack_h2(a, b), glob((uint32_t*)p)); }
  using Word = uint32_t;
  __device__ static Word ld_word(const T* p) { return *glob((const u32_a2*)p); }
  __device__ static void unpack_word(Word v, float, float& a, float& b) { unpack_h2(v, a, b); }
  __device__ static float lo_dev(Word v, float) { return (float)__builtin_bit_cast(f16x2, v).x; }
  __device__ static float hi_dev(Word v, float) { return (float)__builtin_bit_cast(f16x2, v).y; }
  __device__ static float dev_of(T v, float) { return h2f(v); }
};

struct SBF16 {
  using T = uint16_t;
  static constexpr int BYTES = 2;
  __device__ static float ld(const T* p, int64_t e, float w) { return b2f(glob(p)[e]) + w; }
  __device__ static float ld_dev(const T* p, int64_t e, float) { return b2f(glob(p)[e]); }
  __device__ static T enc(float f, float w) { return (T)(pack_b2(f - w, 0.0f) & 0xffffu); }
  __device__ static T enc_dev(float d) { return (T)(pack_b2(d, 0.0f) & 0xffffu); }
  __device__ static void st_dev(T* p, int64_t e, float, float d) { glob(p)[e] = enc_dev(d); }
  template <bool ALIGNED>
  __device__ static void ld2_dev(const T* p, float, float& a, float& b) {
    const uint32_t v = ALIGNED ? *(const uint32_t*)p : *(const u32_a2*)p;
    unpack_b2(v, a, b);
  }
  __device__ static void st2_dev(T* p, float, float a, float b) { __builtin_nontemporal_store(pack_b2(a, b), glob((uint32_t*)p)); }
  using Word = uint32_t;
  __device__ static Word ld_word(const T* p) { return *glob((const u32_a2*)p); }
  __device__ static void unpack_word(Word v, float, float& a, float& b) { unpack_b2(v, a, b); }
  __device__ static float lo_dev(Word v, float) { return __uint_as_float(v << 16); }
  __device__ static float hi_dev(Word v, float) { return __uint_as_float(v & 0xffff0000u); }
  __device__ static float dev_of(T v, float) { return b2f(v); }
};

// ---------------------------------------------------------------------------------------------
// Kernel parameter block: the reference's 96-byte uniform (wgslLbm.ts:6-15) as derived by
// SolverGPU.writeParams (SolverGPU.ts:160-187), plus derived constants (reciprocals of the sponge
// widths, the Smagorinsky factor 18 sqrt2 Cs^2) and fast-division constants.
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
  __host__ __device__ uint32_t div(uint32_t x) const {
    const uint32_t t = (uint32_t)(((uint64_t)m * x) >> 32);
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
  int32_t refX, closure;
  int32_t cpt;        // cells per lane of the bulk update (2: pairs along x, needs even nx; 1)
  int32_t pad1;
  float invSponge;    // 1 / max(nx - spongeStart, 1)
  float invSpongeIn;  // 1 / max(spongeIn, 1)
  float smag;         // 18 sqrt(2) Cs^2
  float inu32;        // 1 / (32 nu0): the learned closure's local-Reynolds input from the wall-distance byte
  FastDiv divx, divy;
};

struct KWheels {
  int32_t count, pad0, pad1, pad2;
  float4 w[32];  // w[2k] = (centre, 1/R), w[2k+1] = (axis, half width), SolverGPU.ts:325-335
};

// One pulled population of a cell next to the body, precomputed from the reference's branch logic
// (inlet, ground + belt, mirrors, periodic z, halfway / Bouzidi walls, rotating wheels with the
// per-cell mass fix; see k_wallprep): d_i = a0 d[e0] + a1 d[e1] + b, with e0 / e1 element indices
// (direction * n + cell) into the population array. A link with one source has e1 = e0, a1 = 0.
struct alignas(16) WallLink {
  int64_t e0, e1;
  float a0, a1, b;
  uint32_t pad;
};
// per wall cell: meta = WM_* bits | x << 8
enum : uint32_t { WM_SOLID = 1u, WM_WALL = 2u, WM_GWALL = 4u };

// Bulk pull sources. A row of cells (fixed y, z) belongs to one of PULL_CLASSES classes,
// cls = yLo | yHi << 1 | zLo << 2 | zHi << 3 (the faces it touches); in a class every direction
// pulls from a fixed source direction and a fixed (dy, dz, dx) shift, which the face rules of the
// pull step decide (wgslLbm.ts:139-170: ground bounce-back, y-z edge mirror, y mirror, z mirror or
// periodic wrap, plain pull). The pull table holds, per class and direction, the element offset of
// the pulled value relative to the pulling cell's row block (row-blocked layout: ((dy + dz ny) Q +
// d) nx + dx). Built on the host (pull_table) for each parameter set.
constexpr int PULL_CLASSES = 16;

template <class L>
__host__ __device__ inline void pull_source(int i, bool yLo, bool yHi, bool zLo, bool zHi, int ground, int sideMode, int nz, int& d,
                                            int& dy, int& dz, int& dx) {
  constexpr bool D3 = L::DIM == 3;
  const int cx = L::c(i, 0), cy = L::c(i, 1), cz = L::c(i, 2);
  const bool yOut = (cy > 0 && yLo) || (cy < 0 && yHi);
  const bool zOut = D3 && ((cz > 0 && zLo) || (cz < 0 && zHi));
  const bool grnd = cy > 0 && yOut && ground != 0;
  const bool yz = yOut && zOut;
  const bool zper = sideMode == 1;
  d = grnd ? opp<L>(i) : yz ? mirYZ<L>(i) : yOut ? mirY<L>(i) : (zOut && !zper) ? mirZ<L>(i) : i;
  dy = (grnd || yOut) ? 0 : -cy;
  // z: unchanged for ground / edge pulls; a y-mirrored pull keeps the plain z shift (it stays in
  // the domain when the z face is not crossed); a crossed z face mirrors (0) or wraps
  dz = (grnd || yz) ? 0 : yOut ? (D3 ? -cz : 0) : zOut ? (zper ? (cz > 0 ? nz - 1 : -(nz - 1)) : 0) : -cz;
  dx = grnd ? 0 : -cx;
}

// Launch layout of one step (see k_step): wall blocks and bulk blocks of 4 waves, in groups of 8
// ordered by the role table.
struct StepArgs {
  const int64_t* pull;     // pull table: PULL_CLASSES x Q element offsets (see pull_source)
  const uint32_t* wlist;   // wall cells
  const uint32_t* wmeta;   // their meta words
  const WallLink* wlink;   // their links, direction-major: wlink[(i - 1) * wcap + t]
  int64_t nwall;
  int64_t wcap;
  uint32_t nwb;     // wall blocks
  uint32_t wpr;     // waves per x-row
  uint32_t nwaves;  // bulk items (one wave-sized chunk of an x-row each, see item_cell)
  uint32_t nbb;     // bulk blocks
  uint32_t xk;      // items per XCD chunk (consecutive items run on one XCD)
  uint32_t pad;
  FastDiv dwpr;     // division by wpr
  FastDiv dxk;      // division by xk
  // role of each group of 8 launch blocks (one block per XCD), in launch order: bit 31 set: wall
  // group (its blocks are wall blocks 8k..8k+7, k in the low bits), else bulk group (bulk blocks
  // 8g..8g+7); (nwb + nbb) / 8 entries
  const uint32_t* role;
};

// The bulk items of a step: items walk the rows in order (x fastest, then y, then z), one wave-sized
// chunk (64 cpt cells) of a row each. Shared by the kernels (item_pos), the host's role table
// (build_roles) and the host check of the layout (tests/native/item_layout_check.hip).
__host__ inline void set_item_layout(StepArgs& A, int64_t nx, int64_t ny, int64_t nz, int cpt) {
  A.wpr = (uint32_t)((nx + 64 * cpt - 1) / (64 * cpt));
  A.dwpr.init(A.wpr);
  A.nwaves = (uint32_t)(ny * nz * A.wpr);
}

// Where lane `lane` of bulk item `item` works: its row (wave-uniform), x0 (the lane's first cell)
// and xs (the first x of the item's chunk, which lanes past the row end load instead of their own).
struct ItemCell {
  uint32_t row;
  int x0, xs;
};
__host__ __device__ inline ItemCell item_cell(const StepArgs& A, uint32_t item, uint32_t lane, int cpt) {
  ItemCell c;
  c.row = A.dwpr.div(item);
  const uint32_t chunk = item - c.row * A.wpr;
  c.xs = (int)(chunk * 64u) * cpt;
  c.x0 = c.xs + (int)lane * cpt;
  return c;
}

// The last item of a row (the one holding its last chunk).
__host__ inline int64_t last_item_of_row(const StepArgs& A, int64_t row) { return row * A.wpr + A.wpr - 1; }

// One domain of a batched step (k_step_batch): its parameters, launch layout, buffers and where its
// blocks sit in the batched grid (each domain's nwb + nbb blocks in a contiguous run, ordered by its
// role table).
struct DomainDesc {
  KParams P;
  StepArgs A;
  const void* f[2];
  const uint8_t* kind;
  uint32_t first;  // first block of the domain (a multiple of 8)
  uint32_t pad;
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

__device__ inline float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// A wave-uniform 64-bit offset kept in scalar registers, so that `pointer argument + uniform
// offset + 32-bit lane offset` selects the SGPR-base / VGPR-offset form of the memory instruction.
__device__ __forceinline__ int64_t uniform_i64(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Inlet equilibrium deviation w_i (3 c_x U + 4.5 c_x^2 U^2 - 1.5 U^2) (rho_in = 1), wgslLbm.ts:146.
template <class L>
__device__ inline float inlet_dev(int i, float U) {
  const float cx = (float)L::c(i, 0);
  return L::w(i) * (3.0f * cx * U + (4.5f * cx * cx - 1.5f) * U * U);
}

// ---------------------------------------------------------------------------------------------
// Collision on deviations (wgslLbm.ts:177-240): moments, non-equilibrium stress, Smagorinsky tau,
// regularized (or BGK) relaxation, outlet sponge and inlet absorbing layer. Moments and the output
// are formed per opposite-direction pair (i, opp i): the even part (mass, stress) and the odd part
// (momentum) once per pair. Algebraically identical to the reference's per-direction form.
// ---------------------------------------------------------------------------------------------
// Sum of c * x over three components with compile-time-known small integer coefficients: zero
// terms are dropped and +-1 terms become adds / subtracts. Accumulation starts from -0.0, the exact
// additive identity, so nothing is left for IEEE rules to keep (0 * x and x + 0.0 are not foldable).
__device__ __forceinline__ float csum(int a, float x, int b, float y, int c, float z) {
  float r = -0.0f;
  if (a == 1) r += x; else if (a == -1) r -= x; else if (a != 0) r += (float)a * x;
  if (b == 1) r += y; else if (b == -1) r -= y; else if (b != 0) r += (float)b * y;
  if (c == 1) r += z; else if (c == -1) r -= z; else if (c != 0) r += (float)c * z;
  return r;
}

// Moments and relaxation parameters of one cell: everything the post-collision populations need
// besides (for BGK) the pre-collision ones. `reg` selects the regularized form.
struct Relax {
  float E;             // even part shared by all directions (mode-dependent, see relax)
  float rho45;         // 4.5 rho
  float ux, uy, uz;    // velocity
  float jx, jy, jz;    // momentum (deviation form: j = rho u)
  float pxx, pyy, pzz, pxy, pxz, pyz;  // Pi^neq
  float k1;            // 4.5 (1 - omega) (regularized)
  float omega, om1;    // 1 / tau, 1 - 1 / tau
  float sIn;           // inlet absorbing-layer weight (0 outside the layer)
  float t0;            // tau before the closure (tau0 + outlet sponge, ground-wall floor)
  float E0, trP;       // E = E0 - k1 trP (regularized) or E0 (BGK)
  float Ex;            // the regularized even part of the directions with c_x != 0 (= E until fold_inlet)
};

// The terms of a cell's update that depend only on its x position: the relaxation time before any
// closure (tau0 + outlet sponge, floored on ground-wall rows) and the inlet absorbing-layer weight.
__device__ __forceinline__ void x_terms(float xpos, bool gwall, const KParams& P, float& t0, float& sIn) {
  const float sp = __builtin_amdgcn_fmed3f((xpos - P.spongeStart) * P.invSponge, 0.0f, 1.0f);
  t0 = P.tau0 + 3.0f * P.spongeNu * sp * sp;
  if (gwall) t0 = fmaxf(t0, P.tauWall);
  const float si = __builtin_amdgcn_fmed3f(1.0f - xpos * P.invSpongeIn, 0.0f, 1.0f);
  sIn = P.spongeIn > 0.0f ? 0.08f * si * si : 0.0f;
}

template <class L>
__device__ __forceinline__ Relax relax(const float (&d)[L::Q], float xpos, bool gwall, bool reg, const KParams& P) {
  constexpr int Q = L::Q;
  constexpr bool D3 = L::DIM == 3;
  Relax m;
  // pair sums / differences
  float s[Q], a[Q];
  float drho = d[0];
#pragma unroll
  for (int i = 1; i < Q; i++) {
    const int o = opp<L>(i);
    if (o < i) continue;
    s[i] = d[i] + d[o];
    a[i] = d[i] - d[o];
    drho += s[i];
  }
  float jx = -0.0f, jy = -0.0f, jz = -0.0f, sxx = -0.0f, syy = -0.0f, szz = -0.0f, sxy = -0.0f, sxz = -0.0f, syz = -0.0f;
#pragma unroll
  for (int i = 1; i < Q; i++) {
    const int o = opp<L>(i);
    if (o < i) continue;
    const int cx = L::c(i, 0), cy = L::c(i, 1), cz = L::c(i, 2);
    if (cx > 0) jx += a[i]; else if (cx < 0) jx -= a[i];
    if (cy > 0) jy += a[i]; else if (cy < 0) jy -= a[i];
    if (cz > 0) jz += a[i]; else if (cz < 0) jz -= a[i];
    if (cx) sxx += s[i];
    if (cy) syy += s[i];
    if (cz) szz += s[i];
    if (cx * cy > 0) sxy += s[i]; else if (cx * cy < 0) sxy -= s[i];
    if (cx * cz > 0) sxz += s[i]; else if (cx * cz < 0) sxz -= s[i];
    if (cy * cz > 0) syz += s[i]; else if (cy * cz < 0) syz -= s[i];
  }
  const float rho = 1.0f + drho;
  const float irho = fast_rcp(rho);
  const float ux = jx * irho, uy = jy * irho, uz = D3 ? jz * irho : 0.0f;
  const float third = 1.0f / 3.0f;
  const float dr3 = drho * third;
  // Pi^neq = sum c c f - rho (u u + delta / 3) = S_dev - j u - drho delta / 3
  const float pxx = sxx - jx * ux - dr3;
  const float pyy = syy - jy * uy - dr3;
  const float pzz = D3 ? szz - jz * uz - dr3 : 0.0f;
  const float pxy = sxy - jx * uy;
  const float pxz = D3 ? sxz - jx * uz : 0.0f;
  const float pyz = D3 ? syz - jy * uz : 0.0f;
  const float usq = D3 ? ux * ux + uy * uy + uz * uz : ux * ux + uy * uy;
  float t0, sIn;
  x_terms(xpos, gwall, P, t0, sIn);
  float tau = t0;
  if (P.smag > 0.0f) {
    const float pn2 = D3 ? pxx * pxx + pyy * pyy + pzz * pzz + 2.0f * (pxy * pxy + pxz * pxz + pyz * pyz)
                         : pxx * pxx + pyy * pyy + 2.0f * (pxy * pxy);
    // hardware square roots (1 ulp): tau only scales the relaxation
    const float piN = __builtin_amdgcn_sqrtf(pn2);
    tau = 0.5f * (t0 + __builtin_amdgcn_sqrtf(t0 * t0 + P.smag * piN * irho));
  }
  const float omega = fast_rcp(tau);
  const float om1 = 1.0f - omega;
  m.sIn = sIn;
  m.rho45 = 4.5f * rho;
  m.ux = ux; m.uy = uy; m.uz = uz;
  m.jx = jx; m.jy = jy; m.jz = jz;
  m.pxx = pxx; m.pyy = pyy; m.pzz = pzz; m.pxy = pxy; m.pxz = pxz; m.pyz = pyz;
  m.omega = omega;
  m.om1 = om1;
  m.t0 = t0;
  m.E0 = drho - 1.5f * rho * usq;
  m.trP = D3 ? (pxx + pyy + pzz) * third : (pxx + pyy) * third;
  if (reg) {
    m.k1 = 4.5f * om1;
    m.E = m.E0 - m.k1 * m.trP;
  } else {
    m.k1 = 0.0f;
    m.E = m.E0;
  }
  m.Ex = m.E;
  return m;
}

// Replaces the relaxation time of a Relax record (the learned closure's tau).
__device__ __forceinline__ void retau(Relax& m, float tau, bool reg) {
  m.omega = fast_rcp(tau);
  m.om1 = 1.0f - m.omega;
  m.k1 = reg ? 4.5f * m.om1 : 0.0f;
  m.E = reg ? m.E0 - m.k1 * m.trP : m.E0;
  m.Ex = m.E;
}

// The inlet absorbing layer folded into a regularized record (wgslLbm.ts:220, :238): the layer's
// mix out + s (feq_i(1, U) - out) = (1 - s) out + s feq_i(1, U) is linear in the record's terms, and
// the inlet equilibrium's deviation w_i (3 c_x U + (4.5 c_x^2 - 1.5) U^2) splits into an even part
// that depends on c_x^2 only and an odd part 3 w_i c_x U. So E, 4.5 rho and k1 scale by 1 - s, the
// even parts of the c_x = 0 and c_x != 0 directions take s (-1.5 U^2) and s (3 U^2) (E and Ex), and
// the momentum becomes (1 - s) j + s (U, 0, 0), once per cell instead of a mix per population. With
// s = 0 (every cell outside the layer) every term keeps its value. Leaves sIn = 0, so the outputs
// apply no mix of their own; the velocity (c.u) is not scaled.
__device__ __forceinline__ void fold_inlet(Relax& m, const KParams& P) {
  const float s = m.sIn, oms = 1.0f - s, U = P.uin;
  const float E = oms * m.E;
  m.E = fmaf(s, -1.5f * U * U, E);
  m.Ex = fmaf(s, 3.0f * U * U, E);
  m.rho45 *= oms;
  m.k1 *= oms;
  m.jx = fmaf(oms, m.jx, s * U);
  m.jy *= oms;
  m.jz *= oms;
  m.sIn = 0.0f;
}

// Inlet absorbing layer: mix toward feq(1, U), wgslLbm.ts:220, :238.
template <class L>
__device__ __forceinline__ float inlet_mix(const Relax& m, int i, float out, const KParams& P) {
  return m.sIn > 0.0f ? out + m.sIn * (inlet_dev<L>(i, P.uin) - out) : out;
}

// Post-collision deviation of the rest direction.
template <class L>
__device__ __forceinline__ float rest_out(const Relax& m, bool reg, float d0, const KParams& P) {
  const float o = reg ? L::w(0) * m.E : m.omega * (L::w(0) * m.E) + m.om1 * d0;
  return inlet_mix<L>(m, 0, o, P);
}

// Post-collision deviations of the pair (i, opp i), i < opp i:
// regularized: w_i (E + 4.5 rho cu^2 + k1 Q_i:Pi) +- 3 w_i c_i.j (Ex in place of E where c_x != 0);
// BGK: omega Deq + (1 - omega) d.
template <class L>
__device__ __forceinline__ void pair_out(const Relax& m, bool reg, int i, float di, float dop, float& oi, float& oo,
                                         const KParams& P) {
  const int o = opp<L>(i);
  const int cx = L::c(i, 0), cy = L::c(i, 1), cz = L::c(i, 2);
  const float cu = csum(cx, m.ux, cy, m.uy, cz, m.uz);
  const float cj = csum(cx, m.jx, cy, m.jy, cz, m.jz);
  const float od = (3.0f * L::w(i)) * cj;
  if (reg) {
    // Q_i : Pi for |c| = 1 (one diagonal entry) or |c|^2 = 2 (two diagonal entries + 2 off-diagonal)
    float qp;
    if (cx && cy) qp = (m.pxx + m.pyy) + (float)(2 * cx * cy) * m.pxy;
    else if (cx && cz) qp = (m.pxx + m.pzz) + (float)(2 * cx * cz) * m.pxz;
    else if (cy && cz) qp = (m.pyy + m.pzz) + (float)(2 * cy * cz) * m.pyz;
    else qp = cx ? m.pxx : (cy ? m.pyy : m.pzz);
    const float ev = L::w(i) * ((cx ? m.Ex : m.E) + m.rho45 * cu * cu + m.k1 * qp);
    oi = ev + od;
    oo = ev - od;
  } else {
    const float ev = L::w(i) * (m.E + m.rho45 * cu * cu);
    oi = (ev + od) + m.om1 * (di - (ev + od));
    oo = (ev - od) + m.om1 * (dop - (ev - od));
  }
  oi = inlet_mix<L>(m, i, oi, P);
  oo = inlet_mix<L>(m, o, oo, P);
}

// BGK collision on deviations in place: f_i - w_i = omega Deq_i + (1 - omega) d_i.
template <class L>
__device__ __forceinline__ void collide_bgk(float (&d)[L::Q], float xpos, bool gwall, const KParams& P) {
  constexpr int Q = L::Q;
  const Relax m = relax<L>(d, xpos, gwall, false, P);
  d[0] = rest_out<L>(m, false, d[0], P);
#pragma unroll
  for (int i = 1; i < Q; i++) {
    const int o = opp<L>(i);
    if (o < i) continue;
    float oi, oo;
    pair_out<L>(m, false, i, d[i], d[o], oi, oo, P);
    d[i] = oi;
    d[o] = oo;
  }
}

// Collision on deviations in place (wall cells use BGK, wgslLbm.ts:186).
template <class L>
__device__ __forceinline__ void collide(float (&d)[L::Q], float xpos, bool wall, bool gwall, const KParams& P) {
  constexpr int Q = L::Q;
  const bool reg = !(wall || P.collision == 1);
  Relax m = relax<L>(d, xpos, gwall, reg, P);
  if (reg) fold_inlet(m, P);
  float out[Q];
  out[0] = rest_out<L>(m, reg, d[0], P);
#pragma unroll
  for (int i = 1; i < Q; i++) {
    const int o = opp<L>(i);
    if (o < i) continue;
    pair_out<L>(m, reg, i, d[i], d[o], out[i], out[o], P);
  }
#pragma unroll
  for (int i = 0; i < Q; i++) d[i] = out[i];
}


}  // namespace wt
