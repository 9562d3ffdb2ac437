// Learned subgrid closure: a per-cell MLP 8-32-32-1 (ReLU hidden layers, softplus output) that
// predicts the eddy viscosity nu_t in place of Smagorinsky (the collision's tau, wgslLbm.ts:214).
// Shared by the step kernel (BF16 MFMA inference), the training kernels (FP32) and the host.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

namespace wt {

constexpr int MLP_IN = 8;
constexpr int MLP_H = 32;
// FP32 master parameters, in this order: W1[32][8] (out x in), b1[32], W2[32][32], b2[32], w3[32], b3
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = 256;
constexpr int MLP_OFF_W2 = 288;
constexpr int MLP_OFF_B2 = 1312;
constexpr int MLP_OFF_W3 = 1344;
constexpr int MLP_OFF_B3 = 1376;
constexpr int MLP_NPARAM = 1377;

// Feature scaling: invariants of degree k are compressed as sign(x) log(1 + |x| / S_REF^k).
constexpr float MLP_S_REF = 1e-3f;
// Targets are clipped at MLP_TARGET_CLIP x nu_scale in training; predictions are capped there too.
constexpr float MLP_TARGET_CLIP = 10.0f;


// Input normalisation and output scale that travel with the weights.
struct MlpNorm {
  float mu[MLP_IN];
  float isig[MLP_IN];  // 1 / sigma
  // the step kernel's form of the normalisation, xn = fma(r, na, nb), on its raw inputs r: base-2
  // logarithms for the compressed inputs (na = ln 2 / sigma), the wall-distance byte (na = 1 / (32
  // sigma)); nb = -mu / sigma
  float na[MLP_IN];
  float nb[MLP_IN];
  float nu_scale;      // nu_t = nu_scale * min(softplus(z), MLP_TARGET_CLIP)
  float smag_floor;    // nu_t >= smag_floor * Smagorinsky's nu_t (the guard)
  float pad[2];
};

// log(1 + y) for y >= 0: the hardware base-2 logarithm on the device, which needs no range
// handling here (its argument is >= 1, never denormal); the inputs only need a few significant
// digits: they are normalised and rounded to BF16 before the network
__host__ __device__ inline float log1p_fast(float y) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_logf(1.0f + y) * 0.69314718f;
#else
  return log1pf(y);
#endif
}

// sign(x) log(1 + |x| / ref), given 1 / ref
__host__ __device__ inline float slog(float x, float iref) {
  const float a = log1p_fast(fabsf(x) * iref);
  return x < 0.0f ? -a : a;
}

// The eight raw (unnormalised) inputs of a cell: six rotation invariants of the strain rate S
// (I1 = tr S, I2 = (I1^2 - tr S^2) / 2, I3 = det S, tr S^2, tr S^3, |S| = sqrt(2 tr S^2)), the wall
// distance d (cells, capped at 4) and the local Reynolds number |u| max(d, 1/2) / nu0 (inu0 = 1 / nu0).
__host__ __device__ inline void mlp_raw_features(float sxx, float syy, float szz, float sxy, float sxz, float syz, float wd,
                                                 float umag, float inu0, float (&x)[MLP_IN]) {
  const float i1 = sxx + syy + szz;
  const float tr2 = sxx * sxx + syy * syy + szz * szz + 2.0f * (sxy * sxy + sxz * sxz + syz * syz);
  const float i2 = 0.5f * (i1 * i1 - tr2);
  const float i3 = sxx * (syy * szz - syz * syz) - sxy * (sxy * szz - syz * sxz) + sxz * (sxy * syz - syy * sxz);
  const float tr3 = i1 * i1 * i1 - 3.0f * i1 * i2 + 3.0f * i3;  // Cayley-Hamilton
#if defined(__HIP_DEVICE_COMPILE__)
  const float m = __builtin_amdgcn_sqrtf(2.0f * tr2);  // hardware square root (about 1 ulp)
#else
  const float m = sqrtf(2.0f * tr2);
#endif
  constexpr float ir = 1.0f / MLP_S_REF, ir2 = ir * ir, ir3 = ir2 * ir;
  x[0] = slog(i1, ir);
  x[1] = slog(i2, ir2);
  x[2] = slog(i3, ir3);
  x[3] = slog(tr2, ir2);
  x[4] = slog(tr3, ir3);
  x[5] = slog(m, ir);
  x[6] = wd;
  x[7] = log1p_fast(umag * fmaxf(wd, 0.5f) * inu0);
}

// The step kernel's inputs, normalised, from the moments of a cell: the same eight inputs as
// mlp_raw_features in a form with fewer instructions. P = -Pi^neq (pxx .. pyz negated: free source
// modifiers), so S = sc P with sc = 3 / (2 rho tau) > 0; an invariant of degree k of S is sc^k times
// that of P, so its compressed form is sign(I(P)) log2(1 + q_k |I(P)|) with q_k = (sc / S_REF)^k,
// the natural logarithm's ln 2 folded into the normalisation (nm.na, nm.nb). wdb: wall distance byte
// (32 per cell); usq: |u|^2; inu32: 1 / (32 nu0).
__device__ __forceinline__ void mlp_step_inputs(float pxx, float pyy, float pzz, float pxy, float pxz, float pyz, float sc, uint32_t wdb,
                                                float usq, float inu32, const MlpNorm& nm, float (&v)[MLP_IN]) {
  const float a = -pxx, b = -pyy, c = -pzz, d = -pxy, e = -pxz, f = -pyz;
  const float i1 = a + b + c;
  const float tr2 = fmaf(a, a, fmaf(b, b, c * c)) + 2.0f * fmaf(d, d, fmaf(e, e, f * f));
  const float i2 = 0.5f * fmaf(i1, i1, -tr2);
  const float i3 = a * fmaf(b, c, -f * f) - d * fmaf(d, c, -f * e) + e * fmaf(d, f, -b * e);
  const float tr3 = fmaf(i1 * i1, i1, 3.0f * fmaf(-i1, i2, i3));
  const float q1 = sc * (1.0f / MLP_S_REF), q2 = q1 * q1, q3 = q2 * q1;
  auto lg = [](float q, float x) { return __builtin_amdgcn_logf(fmaf(q, fabsf(x), 1.0f)); };  // argument >= 1
  v[0] = fmaf(__builtin_copysignf(lg(q1, i1), i1), nm.na[0], nm.nb[0]);
  v[1] = fmaf(__builtin_copysignf(lg(q2, i2), i2), nm.na[1], nm.nb[1]);
  v[2] = fmaf(__builtin_copysignf(lg(q3, i3), i3), nm.na[2], nm.nb[2]);
  v[3] = fmaf(lg(q2, tr2), nm.na[3], nm.nb[3]);
  v[4] = fmaf(__builtin_copysignf(lg(q3, tr3), tr3), nm.na[4], nm.nb[4]);
  v[5] = fmaf(lg(q1, __builtin_amdgcn_sqrtf(2.0f * tr2)), nm.na[5], nm.nb[5]);
  const float wd = (float)wdb;
  v[6] = fmaf(wd, nm.na[6], nm.nb[6]);
  v[7] = fmaf(lg(__builtin_amdgcn_sqrtf(usq) * fmaxf(wd, 16.0f), inu32), nm.na[7], nm.nb[7]);
}

__host__ __device__ inline float softplus(float z) {
#if defined(__HIP_DEVICE_COMPILE__)
  // hardware base-2 exponential and logarithm (1 + e^z >= 1: no range handling needed)
  return z > 20.0f ? z : __builtin_amdgcn_logf(1.0f + __builtin_amdgcn_exp2f(z * 1.44269504f)) * 0.69314718f;
#else
  return z > 20.0f ? z : log1pf(expf(z));
#endif
}

// ---------------------------------------------------------------------------------------------
// Wave-wide BF16 MFMA inference. A wave evaluates 128 cells in four groups of 32, in transposed
// form so that no activation is moved between lanes:
//   H1^T (32 x 32 cells) = W1 (32 x 8) X^T (8 x 32 cells) + b1,  relu   (v_mfma_f32_32x32x8bf16_1k)
//   H2^T (32 x 32 cells) = W2 (32 x 32) H1^T + b2, relu       (two K = 16 chunks, v_mfma_f32_32x32x16_bf16)
//   z^T (32 x 32 cells, all rows equal) = w3 (every row) H2^T  (two K = 16 chunks)
//   nu_t = s min(softplus(z + b3), MLP_TARGET_CLIP).
// Operand layouts (checked on gfx950): A (32 x K): lane l holds row l % 32 and K / 2 consecutive
// k-slots from (K / 2)(l / 32); B (K x 32): lane l holds column l % 32 and the same k-slots; D
// (32 x 32): lane l, register r holds row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32. Register
// block c of D (r = 4c..4c+3) is therefore the B operand of K-chunk c of a K = 8 product, and
// registers 8c..8c+7 the B operand of K-chunk c of a K = 16 product whose k-slot j of lane half h
// stands for hidden unit 16c + 8 (j / 4) + 4h + j % 4 (mlp_hidden_of): the reduction order is
// free, so the A operands are laid out in that order on the host.
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef short bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_hw __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4_hw __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

// hidden unit carried by k-slot j of lane half h in K = 16 chunk c (see the layouts above)
__host__ __device__ constexpr int mlp_hidden_of(int c, int h, int j) { return 16 * c + 8 * (j / 4) + 4 * h + j % 4; }

__device__ inline short to_bf16(float x) {  // round to nearest even (finite inputs)
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// Two floats to packed BF16 (a in the low half) with gfx950's v_cvt_pk_bf16_f32: round to nearest
// even, the same bits as to_bf16 for every non-NaN input (infinities and overflow included).
typedef __bf16 bf16x2_hw __attribute__((ext_vector_type(2)));
typedef float f32x2_hw __attribute__((ext_vector_type(2)));
__device__ inline uint32_t pack_bf16(float a, float b) {
  const f32x2_hw v = {a, b};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_hw));
}

// max(x, 0) as one integer max on the bit pattern: negative floats (sign bit set) are negative
// integers. Gives +0 for -0 and keeps NaNs with a clear sign bit (fmaxf would cost a NaN-quieting
// max per operand besides).
__device__ inline float relu(float x) { return __int_as_float(max(__float_as_int(x), 0)); }

// max(x, 0) on two packed BF16 values with one packed 16-bit integer max (v_pk_max_i16). Rounding
// then clamping gives the same bits as clamping then rounding (a negative value rounds to a
// negative BF16 or to -0, both sent to +0).
typedef short short2_hw __attribute__((ext_vector_type(2)));
__device__ inline uint32_t relu_bf16x2(uint32_t u) {
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(short2_hw, u), (short2_hw){0, 0}));
}

// Device-side network in the layout the MFMA path reads (prepared once from the FP32 master): the
// operands of one lane in one 48-byte record and those of one lane half in one 160-byte record, so
// that two offsets (lane x 48, half x 160) address every read of a group through the instructions'
// immediate offsets. A 48-byte lane stride keeps the 16-byte reads of 16 consecutive lanes on
// distinct LDS banks.
struct alignas(16) MlpLane {
  bf16x8_t w2b[2];  // A operands of layer 2 per K = 16 chunk
  bf16x4_t w1a;     // A operand of layer 1
  uint32_t pad[2];
};
struct alignas(16) MlpHalf {
  float b1d[16];        // layer-1 bias in D layout
  float b2d[16];        // layer-2 bias in D layout
  bf16x8_t w3b[2];      // A operand of the output layer per K = 16 chunk (all 32 rows equal)
};
static_assert(sizeof(MlpLane) == 48 && sizeof(MlpHalf) == 160, "network record strides");
struct alignas(16) MlpDev {
  MlpLane lane[64];
  MlpHalf half[2];
  float b3;
  float pad[3];
  MlpNorm norm;
};

__device__ inline bf16x8_hw relu_pack8(const f32x16_t& acc, int c) {  // D registers 8c..8c+7 -> K = 16 B operand
  u32x4_hw w;
#pragma unroll
  for (int i = 0; i < 4; i++) w[i] = relu_bf16x2(pack_bf16(acc[8 * c + 2 * i], acc[8 * c + 2 * i + 1]));
  return __builtin_bit_cast(bf16x8_hw, w);
}

// Evaluates the network for 64 cells of a wave, one per lane (each lane holds two cells, a and b, and
// calls this once per cell), with no staging of inputs or results in LDS. x: the lane's normalised
// inputs of the cell as packed BF16 pairs (word j holds inputs 2j, 2j + 1). Group 0 of 32 columns
// holds in column j the cell of lane j, group 1 that of lane 32 + j. The K = 8 B operand of a group
// has lane half h hold inputs 4h..4h+3 of its column, so lane j < 32 needs inputs 0-3 and lane 32 + j
// inputs 4-7 of the same cell: one v_permlane32_swap per word (lanes 32-63 of the first operand
// trade places with lanes 0-31 of the second) turns (inputs 0-3, inputs 4-7) of every lane's cell
// into the operands of groups 0 and 1. Every lane receives z of its column in every output register,
// so lane half h takes group h: its own cell. All 64 lanes must call it (MFMA uses the whole wave).
// net: the network, best an LDS copy (stage_net).
__device__ inline bf16x4_t mlp_operand(uint32_t lo, uint32_t hi) {
  typedef uint32_t u32x2_hw __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(bf16x4_t, (u32x2_hw){lo, hi});
}

// One group of 32 cell columns through the network. lo = lane x sizeof(MlpLane), ho = lane half x
// sizeof(MlpHalf): byte offsets of the lane's and the lane half's records. The weights, biases and
// output weights are read from LDS in each group, each just before its product: held in registers
// across the groups (58 of them) they would set the register budget of the step kernel that fuses
// this evaluation; the caller makes the offsets opaque in every group, which keeps the compiler
// from hoisting the reads out of the groups (and leaves no address arithmetic in them).
__device__ inline float mlp_group_z(const MlpDev* __restrict__ net, bf16x4_t xb, int lo, int ho) {
  const MlpLane& ln = *(const MlpLane*)((const char*)net->lane + lo);
  const MlpHalf& hf = *(const MlpHalf*)((const char*)net->half + ho);
  // biases as 16-byte reads (the offsets are opaque, so the reads' alignment comes from the type)
  auto bias = [](const float* b) {
    typedef float f32x4_hw __attribute__((ext_vector_type(4)));
    const f32x4_hw* v = (const f32x4_hw*)b;
    f32x16_t a;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const f32x4_hw t = v[q];
#pragma unroll
      for (int r = 0; r < 4; r++) a[4 * q + r] = t[r];
    }
    return a;
  };
  f32x16_t acc = bias(hf.b1d);
  acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(ln.w1a, xb, acc, 0, 0, 0);
  bf16x8_hw hb0 = relu_pack8(acc, 0), hb1 = relu_pack8(acc, 1);
  acc = bias(hf.b2d);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_hw, ln.w2b[0]), hb0, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_hw, ln.w2b[1]), hb1, acc, 0, 0, 0);
  // output layer on the matrix cores as well: w3 in every row of the A operand, so every register
  // of every lane holds z of the lane's cell column
  hb0 = relu_pack8(acc, 0);
  hb1 = relu_pack8(acc, 1);
  f32x16_t out = {};
  out = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_hw, hf.w3b[0]), hb0, out, 0, 0, 0);
  out = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_hw, hf.w3b[1]), hb1, out, 0, 0, 0);
  return out[0];
}

// The network's output z + b3 for one cell of every lane (x: the cell's packed inputs). A lane's two
// cells go through two calls, cell a first, so that only one cell's operands are live at a time.
__device__ inline float mlp_wave_z(const MlpDev* __restrict__ net, const uint32_t (&x)[4]) {
  const int l = threadIdx.x & 63;
  const auto s0 = __builtin_amdgcn_permlane32_swap(x[0], x[2], false, false);
  const auto s1 = __builtin_amdgcn_permlane32_swap(x[1], x[3], false, false);
  int lo = l * (int)sizeof(MlpLane), ho = (l >> 5) * (int)sizeof(MlpHalf);
  // the wave's issue priority is raised over its dependent matrix chains, so that each chain moves
  // on as soon as its operands are ready while the other waves' vector work fills the gaps
  // (0.4 % faster at 256x128x128, five waves per SIMD)
  __builtin_amdgcn_s_setprio(3);
  // one group at a time (not interleaved by the compiler, which would hold two groups' accumulators)
  float z = 0.0f;
#pragma unroll 1
  for (int g = 0; g < 2; g++) {  // g = 0: the cells of lanes 0-31, g = 1: those of lanes 32-63
    asm volatile("" : "+v"(lo), "+v"(ho));
    const float zg = mlp_group_z(net, mlp_operand(g ? s0[1] : s0[0], g ? s1[1] : s1[0]), lo, ho);
    if ((l >= 32) == (g == 1)) z = zg;
  }
  __builtin_amdgcn_s_setprio(0);
  return z + net->b3;
}

__device__ inline float mlp_nu_of(const MlpDev* __restrict__ net, float z) {
  return net->norm.nu_scale * fminf(softplus(z), MLP_TARGET_CLIP);
}

// Block-wide copy of the network into LDS, in two halves so that a kernel can issue its own global
// loads between them: stage_net_load reads one 16-byte piece per thread (the network fits in one
// piece per thread of a 256-thread block), stage_net_store writes it to LDS and synchronises the block
// (every thread of the block must call both). The vector-memory wait counter drains in issue order:
// the LDS write waits for the piece only, and loads issued after it stay in flight.
constexpr int MLP_DEV_PIECES = (int)(sizeof(MlpDev) / 16);
static_assert(sizeof(MlpDev) % 16 == 0 && MLP_DEV_PIECES <= 256, "the network is staged as one 16-byte piece per thread");
__device__ inline u32x4_hw stage_net_load(const MlpDev* __restrict__ src) {
  const u32x4_hw* a = (const u32x4_hw*)src;
  return threadIdx.x < (unsigned)MLP_DEV_PIECES ? a[threadIdx.x] : u32x4_hw{0u, 0u, 0u, 0u};
}
__device__ inline void stage_net_store(MlpDev* dst, u32x4_hw v) {
  if (threadIdx.x < (unsigned)MLP_DEV_PIECES) ((u32x4_hw*)dst)[threadIdx.x] = v;
  __syncthreads();
}
__device__ inline void stage_net(const MlpDev* __restrict__ src, MlpDev* dst) { stage_net_store(dst, stage_net_load(src)); }

}  // namespace wt
