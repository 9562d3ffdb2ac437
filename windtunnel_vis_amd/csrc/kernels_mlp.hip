// Learned subgrid closure: wall distance, training data from a fine run (2x box filter), the
// network's training kernels (FP32 forward / backward / Adam) and a stand-alone inference kernel on
// the MFMA path the step kernel uses.
#include <hip/hip_runtime.h>
#include <algorithm>

#include "lbm.hpp"
#include "launch.hpp"
#include "mlp.hpp"

namespace wt {

namespace {
constexpr int BLOCK = 256;
inline unsigned blocks_for(int64_t n, int b = BLOCK) { return (unsigned)((n + b - 1) / b); }

__device__ inline void cell_coords(int64_t c, int nx, int ny, int& x, int& y, int& z) {
  x = (int)(c % nx);
  const int64_t r = c / nx;
  y = (int)(r % ny);
  z = (int)(r / ny);
}
}  // namespace

// ---------------------------------------------------------------------------------------------
// Wall distance: for every cell, the distance between its centre and the nearest solid cell
// centre, capped at 4 cells (0 on solid cells). Only surface solid cells (a non-solid 6-neighbour)
// can be nearest, so they scatter squared distances over their radius-4 ball (atomicMin).
// Stored as a byte: round(32 d).
// ---------------------------------------------------------------------------------------------
constexpr int WD_R = 4;

__global__ __launch_bounds__(BLOCK) void k_wdist_init(const uint32_t* __restrict__ flags, uint32_t* __restrict__ d2, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) d2[i] = flags[i] != 0u ? 0u : (uint32_t)(WD_R * WD_R);
}

__global__ __launch_bounds__(BLOCK) void k_wdist_scatter(const uint32_t* __restrict__ flags, uint32_t* __restrict__ d2, int nx, int ny,
                                                         int nz) {
  const int64_t n = (int64_t)nx * ny * nz;
  const int64_t c = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (c >= n || flags[c] == 0u) return;
  int x, y, z;
  cell_coords(c, nx, ny, x, y, z);
  auto solid = [&](int a, int b, int e) -> bool {
    if (a < 0 || b < 0 || e < 0 || a >= nx || b >= ny || e >= nz) return true;  // outside: not a fluid neighbour
    return flags[(int64_t)a + (int64_t)nx * ((int64_t)b + (int64_t)ny * e)] != 0u;
  };
  if (solid(x + 1, y, z) && solid(x - 1, y, z) && solid(x, y + 1, z) && solid(x, y - 1, z) && solid(x, y, z + 1) && solid(x, y, z - 1))
    return;
  for (int dz = -WD_R; dz <= WD_R; dz++) {
    const int zz = z + dz;
    if (zz < 0 || zz >= nz) continue;
    for (int dy = -WD_R; dy <= WD_R; dy++) {
      const int yy = y + dy;
      if (yy < 0 || yy >= ny) continue;
      for (int dx = -WD_R; dx <= WD_R; dx++) {
        const int xx = x + dx;
        const uint32_t q = (uint32_t)(dx * dx + dy * dy + dz * dz);
        if (xx < 0 || xx >= nx || q >= (uint32_t)(WD_R * WD_R)) continue;
        atomicMin(&d2[(int64_t)xx + (int64_t)nx * ((int64_t)yy + (int64_t)ny * zz)], q);
      }
    }
  }
}

__global__ __launch_bounds__(BLOCK) void k_wdist_pack(const uint32_t* __restrict__ d2, uint8_t* __restrict__ wd, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) wd[i] = (uint8_t)__float2int_rn(__builtin_sqrtf((float)d2[i]) * 32.0f);
}

void launch_wall_distance(hipStream_t st, const uint32_t* flags, uint32_t* tmp, uint8_t* out, int nx, int ny, int nz) {
  const int64_t n = (int64_t)nx * ny * nz;
  hipLaunchKernelGGL(k_wdist_init, dim3(blocks_for(n)), dim3(BLOCK), 0, st, flags, tmp, n);
  hipLaunchKernelGGL(k_wdist_scatter, dim3(blocks_for(n)), dim3(BLOCK), 0, st, flags, tmp, nx, ny, nz);
  hipLaunchKernelGGL(k_wdist_pack, dim3(blocks_for(n)), dim3(BLOCK), 0, st, tmp, out, n);
}

// ---------------------------------------------------------------------------------------------
// Training data from a fine run. Pass 1 (per coarse cell = 2x2x2 fine cells): the block mean
// velocity and the subgrid stress tau_ij = <(u_i - <u_i>)(u_j - <u_j>)> (two-pass, no
// cancellation); a block touching a solid fine cell is marked bad. Pass 2 (per coarse cell): the
// filtered strain by central differences of the block means and the least-squares numerator
// -tau_dev : S and denominator 2 S : S. Pass 3: the target nu_t = max(0, sum num / sum den) over
// the 3x3x3 neighbourhood (coarse lattice units) and the features of the coarse cell; cells on
// the domain faces or next to a bad block are skipped. Samples are written in cell
// order (per-block counts, host scan, ballot ranks), so a dataset is reproducible.
// Record: 8 raw features, target, coarse cell index (as bits).
// ---------------------------------------------------------------------------------------------
constexpr int COARSE_REC = 10;

template <class S>
__global__ __launch_bounds__(BLOCK) void k_coarse_fields(const typename S::T* __restrict__ f, const uint32_t* __restrict__ flags,
                                                         int fnx, int fny, int fnz, float* __restrict__ cf, uint8_t* __restrict__ bad) {
  using L = D3Q19;
  const int cnx = fnx / 2, cny = fny / 2, cnz = fnz / 2;
  const int64_t nc = (int64_t)cnx * cny * cnz;
  const int64_t c = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (c >= nc) return;
  int x, y, z;
  cell_coords(c, cnx, cny, x, y, z);
  const int64_t nf = (int64_t)fnx * fny * fnz;
  float u[8][3];
  bool b = false;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int64_t frow = (int64_t)(2 * y + ((k >> 1) & 1)) + (int64_t)fny * (2 * z + (k >> 2));
    const int fx = 2 * x + (k & 1);
    const int64_t e = (int64_t)fx + (int64_t)fnx * frow;
    b |= flags[e] != 0u;
    float dr = 0.f, jx = 0.f, jy = 0.f, jz = 0.f;
#pragma unroll
    for (int i = 0; i < L::Q; i++) {
      const float d = S::ld_dev(f, pop_at(frow, i, fx, L::Q, fnx), L::w(i));
      dr += d;
      if (L::c(i, 0)) jx += L::c(i, 0) * d;
      if (L::c(i, 1)) jy += L::c(i, 1) * d;
      if (L::c(i, 2)) jz += L::c(i, 2) * d;
    }
    const float rho = 1.0f + dr;
    u[k][0] = jx / rho;
    u[k][1] = jy / rho;
    u[k][2] = jz / rho;
  }
  float m[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < 8; k++)
#pragma unroll
    for (int a = 0; a < 3; a++) m[a] += u[k][a];
#pragma unroll
  for (int a = 0; a < 3; a++) m[a] *= 0.125f;
  float t[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};  // xx yy zz xy xz yz
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const float ax = u[k][0] - m[0], ay = u[k][1] - m[1], az = u[k][2] - m[2];
    t[0] += ax * ax; t[1] += ay * ay; t[2] += az * az;
    t[3] += ax * ay; t[4] += ax * az; t[5] += ay * az;
  }
  float* o = cf + c * 9;
  o[0] = m[0]; o[1] = m[1]; o[2] = m[2];
#pragma unroll
  for (int a = 0; a < 6; a++) o[3 + a] = t[a] * 0.125f;
  bad[c] = b ? 1 : 0;
}

// Filtered strain of coarse cell c from the block means: central differences, one-sided along an
// axis where one neighbour block holds a solid cell (so cells beside the body get a strain and a
// training target too); false on the domain faces, for a bad block, and where both neighbours
// along an axis are bad.
__device__ inline bool coarse_strain(const float* __restrict__ cf, const uint8_t* __restrict__ bad, int cnx, int cny, int cnz, int64_t c,
                                     float (&S)[6]) {
  int x, y, z;
  cell_coords(c, cnx, cny, x, y, z);
  if (x < 1 || y < 1 || z < 1 || x >= cnx - 1 || y >= cny - 1 || z >= cnz - 1) return false;
  if (bad[c]) return false;
  const int64_t sx = 1, sy = cnx, sz = (int64_t)cnx * cny;
  float g[3][3];  // g[i][j] = d u_j / d x_i
  const int64_t st[3] = {sx, sy, sz};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const bool up = !bad[c + st[i]], dn = !bad[c - st[i]];
    if (!up && !dn) return false;
    const int64_t a = up ? c + st[i] : c, b = dn ? c - st[i] : c;
    const float h = (up && dn) ? 0.5f : 1.0f;
#pragma unroll
    for (int j = 0; j < 3; j++) g[i][j] = h * (cf[a * 9 + j] - cf[b * 9 + j]);
  }
  S[0] = g[0][0]; S[1] = g[1][1]; S[2] = g[2][2];
  S[3] = 0.5f * (g[0][1] + g[1][0]); S[4] = 0.5f * (g[0][2] + g[2][0]); S[5] = 0.5f * (g[1][2] + g[2][1]);
  return true;
}

// Per coarse cell: the least-squares numerator -tau_dev : S and denominator 2 S : S of the eddy
// viscosity (both 0 where the strain is not defined), for the neighbourhood average of the target.
__global__ __launch_bounds__(BLOCK) void k_coarse_nd(const float* __restrict__ cf, const uint8_t* __restrict__ bad, int cnx, int cny,
                                                     int cnz, float2* __restrict__ nd) {
  const int64_t nc = (int64_t)cnx * cny * cnz;
  const int64_t c = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (c >= nc) return;
  float S[6];
  float2 r = make_float2(0.0f, 0.0f);
  if (coarse_strain(cf, bad, cnx, cny, cnz, c, S)) {
    const float* t = cf + c * 9 + 3;
    const float tr3 = (t[0] + t[1] + t[2]) * (1.0f / 3.0f);
    r.x = -((t[0] - tr3) * S[0] + (t[1] - tr3) * S[1] + (t[2] - tr3) * S[2] + 2.0f * (t[3] * S[3] + t[4] * S[4] + t[5] * S[5]));
    r.y = 2.0f * (S[0] * S[0] + S[1] * S[1] + S[2] * S[2] + 2.0f * (S[3] * S[3] + S[4] * S[4] + S[5] * S[5]));
  }
  nd[c] = r;
}

// Record of coarse cell c: features from its filtered strain, wall distance and speed; target
// nu_t = max(0, sum num / sum den) over the 3x3x3 neighbourhood (the least-squares eddy viscosity
// of the neighbourhood, Germano-style averaging); false if the cell is not usable.
__device__ inline bool coarse_sample(const float* __restrict__ cf, const uint8_t* __restrict__ bad, const float2* __restrict__ nd,
                                     const uint8_t* __restrict__ wd, int cnx, int cny, int cnz, float nu0, int64_t c,
                                     float (&rec)[COARSE_REC]) {
  float S[6];
  if (!coarse_strain(cf, bad, cnx, cny, cnz, c, S)) return false;
  float num = 0.0f, den = 0.0f;
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        const float2 v = nd[c + dx + (int64_t)cnx * (dy + (int64_t)cny * dz)];
        num += v.x;
        den += v.y;
      }
  const float nut = fmaxf(0.0f, num / (den + 1e-30f));
  const float* ub = cf + c * 9;
  const float um = sqrtf(ub[0] * ub[0] + ub[1] * ub[1] + ub[2] * ub[2]);
  float xf[MLP_IN];
  mlp_raw_features(S[0], S[1], S[2], S[3], S[4], S[5], (float)wd[c] * (1.0f / 32.0f), um, 1.0f / nu0, xf);
#pragma unroll
  for (int k = 0; k < MLP_IN; k++) rec[k] = xf[k];
  rec[8] = nut;
  rec[9] = __uint_as_float((uint32_t)c);
  return true;
}

__global__ __launch_bounds__(BLOCK) void k_coarse_count(const float* __restrict__ cf, const uint8_t* __restrict__ bad,
                                                        const float2* __restrict__ nd, const uint8_t* __restrict__ wd, int cnx, int cny,
                                                        int cnz, float nu0, uint32_t* __restrict__ counts) {
  __shared__ uint32_t s_n;
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  const int64_t nc = (int64_t)cnx * cny * cnz;
  const int64_t c = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  float rec[COARSE_REC];
  const bool ok = c < nc && coarse_sample(cf, bad, nd, wd, cnx, cny, cnz, nu0, c, rec);
  if (ok) atomicAdd(&s_n, 1u);
  __syncthreads();
  if (threadIdx.x == 0) counts[blockIdx.x] = s_n;
}

__global__ __launch_bounds__(BLOCK) void k_coarse_samples(const float* __restrict__ cf, const uint8_t* __restrict__ bad,
                                                          const float2* __restrict__ nd, const uint8_t* __restrict__ wd, int cnx,
                                                          int cny, int cnz, float nu0, const uint64_t* __restrict__ offsets,
                                                          float* __restrict__ out) {
  __shared__ uint32_t s_wave[BLOCK / 64];
  const int64_t nc = (int64_t)cnx * cny * cnz;
  const int64_t c = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  float rec[COARSE_REC];
  const bool ok = c < nc && coarse_sample(cf, bad, nd, wd, cnx, cny, cnz, nu0, c, rec);
  const uint64_t mask = __ballot(ok);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) s_wave[w] = (uint32_t)__popcll(mask);
  __syncthreads();
  uint32_t base = 0;
  for (int k = 0; k < w; k++) base += s_wave[k];
  if (!ok) return;
  const uint32_t rank = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
  float* o = out + (offsets[blockIdx.x] + rank) * COARSE_REC;
#pragma unroll
  for (int k = 0; k < COARSE_REC; k++) o[k] = rec[k];
}

int64_t coarse_blocks(int64_t nc) { return (int64_t)blocks_for(nc); }

void launch_coarse_fields(hipStream_t st, int prec, const void* f, const uint32_t* flags, int fnx, int fny, int fnz, float* cf,
                          uint8_t* bad) {
  const int64_t nc = (int64_t)(fnx / 2) * (fny / 2) * (fnz / 2);
  const dim3 g(blocks_for(nc));
  if (prec == 0) hipLaunchKernelGGL((k_coarse_fields<SF32>), g, dim3(BLOCK), 0, st, (const float*)f, flags, fnx, fny, fnz, cf, bad);
  else if (prec == 1)
    hipLaunchKernelGGL((k_coarse_fields<SF16>), g, dim3(BLOCK), 0, st, (const uint16_t*)f, flags, fnx, fny, fnz, cf, bad);
  else hipLaunchKernelGGL((k_coarse_fields<SBF16>), g, dim3(BLOCK), 0, st, (const uint16_t*)f, flags, fnx, fny, fnz, cf, bad);
}

void launch_coarse_count(hipStream_t st, const float* cf, const uint8_t* bad, float2* nd, const uint8_t* wd, int cnx, int cny, int cnz,
                         float nu0, uint32_t* counts) {
  const int64_t nc = (int64_t)cnx * cny * cnz;
  hipLaunchKernelGGL(k_coarse_nd, dim3(blocks_for(nc)), dim3(BLOCK), 0, st, cf, bad, cnx, cny, cnz, nd);
  hipLaunchKernelGGL(k_coarse_count, dim3(blocks_for(nc)), dim3(BLOCK), 0, st, cf, bad, nd, wd, cnx, cny, cnz, nu0, counts);
}

void launch_coarse_samples(hipStream_t st, const float* cf, const uint8_t* bad, const float2* nd, const uint8_t* wd, int cnx, int cny,
                           int cnz, float nu0, const uint64_t* offsets, float* out) {
  const int64_t nc = (int64_t)cnx * cny * cnz;
  hipLaunchKernelGGL(k_coarse_samples, dim3(blocks_for(nc)), dim3(BLOCK), 0, st, cf, bad, nd, wd, cnx, cny, cnz, nu0, offsets, out);
}

// ---------------------------------------------------------------------------------------------
// Training (FP32). k_mlp_grad: a block of 128 threads takes 128 samples (by a permutation):
// forward and backward per thread, activations and deltas in LDS, then each thread sums a slice
// of the 1,377 parameter gradients over the block's samples -> per-block partials (+ loss).
// k_grad_reduce sums the partials in a fixed order; k_adam updates the FP32 master parameters.
// ---------------------------------------------------------------------------------------------
constexpr int TB = 128;
constexpr int GP = MLP_NPARAM + 1;  // partial record: gradients + loss

__global__ __launch_bounds__(TB) void k_mlp_grad(const float* __restrict__ data, const uint32_t* __restrict__ perm, int64_t first,
                                                 int64_t count, float inv_batch, const float* __restrict__ params, MlpNorm norm,
                                                 float* __restrict__ partials) {
  __shared__ float P[MLP_NPARAM];
  __shared__ float X[TB][MLP_IN + 1];
  __shared__ float H1[TB][MLP_H + 1];
  __shared__ float H2[TB][MLP_H + 1];
  __shared__ float D1[TB][MLP_H + 1];
  __shared__ float D2[TB][MLP_H + 1];
  __shared__ float DZ[TB];
  __shared__ float LS[TB];
  const int t = threadIdx.x;
  for (int e = t; e < MLP_NPARAM; e += TB) P[e] = params[e];
  __syncthreads();
  const int64_t s = (int64_t)blockIdx.x * TB + t;
  const float* W1 = P + MLP_OFF_W1;
  const float* B1 = P + MLP_OFF_B1;
  const float* W2 = P + MLP_OFF_W2;
  const float* B2 = P + MLP_OFF_B2;
  const float* W3 = P + MLP_OFF_W3;
  if (s < count) {
    const float* rec = data + (int64_t)perm[first + s] * COARSE_REC;
    float x[MLP_IN];
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) {
      x[i] = (rec[i] - norm.mu[i]) * norm.isig[i];
      X[t][i] = x[i];
    }
    const float target = fminf(rec[8] / norm.nu_scale, MLP_TARGET_CLIP);
    float h1[MLP_H];
#pragma unroll
    for (int o = 0; o < MLP_H; o++) {
      float a = B1[o];
#pragma unroll
      for (int i = 0; i < MLP_IN; i++) a = fmaf(W1[o * MLP_IN + i], x[i], a);
      h1[o] = fmaxf(a, 0.0f);
      H1[t][o] = h1[o];
    }
    float z = P[MLP_OFF_B3];
    for (int o = 0; o < MLP_H; o++) {
      float a = B2[o];
#pragma unroll
      for (int i = 0; i < MLP_H; i++) a = fmaf(W2[o * MLP_H + i], h1[i], a);
      const float h2 = fmaxf(a, 0.0f);
      H2[t][o] = h2;
      z = fmaf(W3[o], h2, z);
    }
    const float pred = softplus(z);
    const float r = pred - target;
    LS[t] = r * r;
    const float dz = 2.0f * r * inv_batch / (1.0f + expf(-z));
    DZ[t] = dz;
    for (int o = 0; o < MLP_H; o++) D2[t][o] = H2[t][o] > 0.0f ? dz * W3[o] : 0.0f;
    for (int i = 0; i < MLP_H; i++) {
      float a = 0.0f;
      for (int o = 0; o < MLP_H; o++) a = fmaf(W2[o * MLP_H + i], D2[t][o], a);
      D1[t][i] = h1[i] > 0.0f ? a : 0.0f;
    }
  } else {
    for (int i = 0; i < MLP_IN; i++) X[t][i] = 0.0f;
    for (int o = 0; o < MLP_H; o++) H1[t][o] = H2[t][o] = D1[t][o] = D2[t][o] = 0.0f;
    DZ[t] = 0.0f;
    LS[t] = 0.0f;
  }
  __syncthreads();
  float* out = partials + (int64_t)blockIdx.x * GP;
  for (int e = t; e < GP; e += TB) {
    float a = 0.0f;
    if (e < MLP_OFF_B1) {
      const int o = e / MLP_IN, i = e % MLP_IN;
      for (int k = 0; k < TB; k++) a = fmaf(D1[k][o], X[k][i], a);
    } else if (e < MLP_OFF_W2) {
      const int o = e - MLP_OFF_B1;
      for (int k = 0; k < TB; k++) a += D1[k][o];
    } else if (e < MLP_OFF_B2) {
      const int o = (e - MLP_OFF_W2) / MLP_H, i = (e - MLP_OFF_W2) % MLP_H;
      for (int k = 0; k < TB; k++) a = fmaf(D2[k][o], H1[k][i], a);
    } else if (e < MLP_OFF_W3) {
      const int o = e - MLP_OFF_B2;
      for (int k = 0; k < TB; k++) a += D2[k][o];
    } else if (e < MLP_OFF_B3) {
      const int i = e - MLP_OFF_W3;
      for (int k = 0; k < TB; k++) a = fmaf(DZ[k], H2[k][i], a);
    } else if (e == MLP_OFF_B3) {
      for (int k = 0; k < TB; k++) a += DZ[k];
    } else {
      for (int k = 0; k < TB; k++) a += LS[k];
    }
    out[e] = a;
  }
}

// grad[e] = sum over blocks (fixed order, FP64); loss (sum of squared residuals) in lossacc.
__global__ __launch_bounds__(BLOCK) void k_grad_reduce(const float* __restrict__ partials, int nblocks, float* __restrict__ grad,
                                                       double* __restrict__ lossacc) {
  const int e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= GP) return;
  double a = 0.0;
  for (int b = 0; b < nblocks; b++) a += partials[(int64_t)b * GP + e];
  if (e < MLP_NPARAM) grad[e] = (float)a;
  else *lossacc += a;
}

__global__ __launch_bounds__(BLOCK) void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                                                const float* __restrict__ g, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= MLP_NPARAM) return;
  const float gg = g[e];
  const float mm = b1 * m[e] + (1.0f - b1) * gg;
  const float vv = b2 * v[e] + (1.0f - b2) * gg * gg;
  m[e] = mm;
  v[e] = vv;
  p[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

int64_t grad_partial_floats(int64_t samples) { return ((samples + TB - 1) / TB) * GP; }

void launch_mlp_grad(hipStream_t st, const float* data, const uint32_t* perm, int64_t first, int64_t count, float inv_batch,
                     const float* params, const MlpNorm& norm, float* partials, float* grad, double* lossacc) {
  const int nb = (int)((count + TB - 1) / TB);
  hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(TB), 0, st, data, perm, first, count, inv_batch, params, norm, partials);
  hipLaunchKernelGGL(k_grad_reduce, dim3(blocks_for(GP)), dim3(BLOCK), 0, st, partials, nb, grad, lossacc);
}

void launch_adam(hipStream_t st, float* p, float* m, float* v, const float* g, float lr, float b1, float b2, float eps, int64_t t) {
  const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(MLP_NPARAM)), dim3(BLOCK), 0, st, p, m, v, g, lr, b1, b2, eps, c1, c2);
}

// ---------------------------------------------------------------------------------------------
// Stand-alone inference on the step kernel's MFMA path: raw features (n x 8) -> nu_t (n).
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_mlp_infer(const MlpDev* __restrict__ gnet, const float* __restrict__ raw, int64_t n,
                                                     float* __restrict__ nu) {
  __shared__ MlpDev snet;
  stage_net(gnet, &snet);
  const MlpDev* net = &snet;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t base = ((int64_t)blockIdx.x * (BLOCK / 64) + w) * 128;
  // every lane evaluates two cells (2l, 2l + 1), as the step kernel's lanes do
  uint32_t x[2][MLP_IN / 2];
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int64_t c = base + 2 * l + h;
    float v[MLP_IN];
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) v[i] = c < n ? (raw[c * MLP_IN + i] - net->norm.mu[i]) * net->norm.isig[i] : 0.0f;
#pragma unroll
    for (int i = 0; i < MLP_IN / 2; i++) x[h][i] = (uint32_t)(uint16_t)to_bf16(v[2 * i]) | ((uint32_t)(uint16_t)to_bf16(v[2 * i + 1]) << 16);
  }
  float r[2];
#pragma unroll
  for (int h = 0; h < 2; h++) r[h] = mlp_nu_of(net, mlp_wave_z(net, x[h]));
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int64_t c = base + 2 * l + h;
    if (c < n) nu[c] = r[h];
  }
}

void launch_mlp_infer(hipStream_t st, const MlpDev* net, const float* raw, int64_t n, float* nu) {
  const int64_t waves = (n + 127) / 128;
  hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((waves + 3) / 4)), dim3(BLOCK), 0, st, net, raw, n, nu);
}

}  // namespace wt
