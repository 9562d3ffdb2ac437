// CPU oracle for the lattice-Boltzmann wind-tunnel hot path.
//
// TEST INFRASTRUCTURE ONLY. This file restates, on the CPU and in FP64 (or FP32 for the timed CPU
// baseline), the semantics of the reference's WGSL kernels so that the HIP kernels can be checked
// against it. Nothing in the product library links or calls it; only the tests, the smoke check and
// the CPU-baseline leg of the benchmark load it.
//
// What it restates (paths relative to the reference repository):
//   lattice tables            src/solver/lattice.ts:20-60
//   parameter derivation      src/solver/SolverGPU.ts:160-187
//   init kernel               src/solver/wgslLbm.ts:296-341
//   fused stream + collide    src/solver/wgslLbm.ts:134-293 (pull branches :139-175, contact patch
//                             :276-287, wheel mass fix :289-291, collision :177-240)
//   sdf encoding / link q     src/solver/SolverGPU.ts:312-322, src/solver/wgslLbm.ts:31-43
//   wheel velocity            src/solver/wgslLbm.ts:21-28, SolverGPU.ts:325-335
//   macro + momentum exchange src/solver/wgslLbm.ts:353-489
//   final reduction           src/solver/wgslLbm.ts:491-521
//
// Storage emulation: populations are kept in T; after every store they can be rounded the way the
// device stores them (FP32 values; FP16 or BF16 deviations f - w_i, round-to-nearest-even).
//
// Oracle-only extension: `periodic_x` wraps x instead of inlet/outlet (for the shear-wave test).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

struct or_params {
  int32_t nx, ny, nz;
  double u, nu, cs;
  int32_t ground;      // 0 freeslip, 1 noslip, 2 moving belt
  int32_t sides;       // 0 freeslip, 1 periodic (z)
  double sponge_nu, sponge_start, sponge_in, tau_wall;
  int32_t interp;      // 1 Bouzidi, 0 halfway bounce-back
  int32_t collision;   // 0 regularized, 1 BGK
  double ema_alpha;
  int32_t storage;     // 0 none (full T), 1 f32, 2 f16 deviation, 3 bf16 deviation
  int32_t periodic_x;  // oracle-only
  int32_t threads;     // OpenMP threads (0 = default)
};

// The learned closure (no reference counterpart: it replaces the Smagorinsky tau of
// wgslLbm.ts:214), as the product defines it (include/wt.h, wt_mlp): an 8-32-32-1 network on six
// strain invariants, the wall distance and the local Reynolds number, BF16 inputs, weights and
// first hidden layer, softplus output capped at 10 x nu_scale, guarded by smag_floor x the
// Smagorinsky eddy viscosity. It runs on bulk cells: fluid cells whose group of `pair` cells along
// x (the product's cells per lane) holds no cell with a plain pull from a solid and no solid cell;
// the other cells keep Smagorinsky.
struct or_mlp {
  double params[1377];  // W1[32][8], b1[32], W2[32][32], b2[32], w3[32], b3
  double mu[8], sigma[8];
  double nu_scale, smag_floor;
  int32_t pair;          // 1 or 2 (2 needs an even nx)
  int32_t pad;
};

}  // extern "C"

namespace {

struct Lattice {
  int q = 0, dim = 0;
  int c[19][3];
  double w[19];
  int opp[19], mY[19], mZ[19], mYZ[19];
};

// lattice.ts:39-60 velocity sets, with opp / mirrors found by search as lattice.ts:20-37 does.
Lattice make_lattice(bool d3) {
  Lattice L;
  static const int c9[9][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {-1, 0, 0}, {0, -1, 0},
                               {1, 1, 0}, {-1, 1, 0}, {-1, -1, 0}, {1, -1, 0}};
  static const int c19[19][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},
                                 {0, 0, 1},  {0, 0, -1}, {1, 1, 0},  {-1, -1, 0}, {1, 0, 1},
                                 {-1, 0, -1}, {0, 1, 1}, {0, -1, -1}, {1, -1, 0}, {-1, 1, 0},
                                 {1, 0, -1}, {-1, 0, 1}, {0, 1, -1}, {0, -1, 1}};
  if (d3) {
    L.q = 19; L.dim = 3;
    for (int i = 0; i < 19; i++) {
      for (int a = 0; a < 3; a++) L.c[i][a] = c19[i][a];
      L.w[i] = i == 0 ? 1.0 / 3.0 : (i < 7 ? 1.0 / 18.0 : 1.0 / 36.0);
    }
  } else {
    L.q = 9; L.dim = 2;
    for (int i = 0; i < 9; i++) {
      for (int a = 0; a < 3; a++) L.c[i][a] = c9[i][a];
      L.w[i] = i == 0 ? 4.0 / 9.0 : (i < 5 ? 1.0 / 9.0 : 1.0 / 36.0);
    }
  }
  auto find = [&](int x, int y, int z) {
    for (int k = 0; k < L.q; k++)
      if (L.c[k][0] == x && L.c[k][1] == y && L.c[k][2] == z) return k;
    return -1;
  };
  for (int i = 0; i < L.q; i++) {
    L.opp[i] = find(-L.c[i][0], -L.c[i][1], -L.c[i][2]);
    L.mY[i] = find(L.c[i][0], -L.c[i][1], L.c[i][2]);
    L.mZ[i] = find(L.c[i][0], L.c[i][1], -L.c[i][2]);
    L.mYZ[i] = find(L.c[i][0], -L.c[i][1], -L.c[i][2]);
  }
  return L;
}

// IEEE binary16 round-to-nearest-even of a double, returned as the double it represents.
double round_half(double x) {
  if (!std::isfinite(x)) return x;
  double a = std::fabs(x);
  if (a >= 65520.0) return std::copysign(INFINITY, x);
  int e;
  std::frexp(a, &e);                // a = m * 2^e, m in [0.5, 1)
  int exp = e - 1;                  // a in [2^exp, 2^(exp+1))
  if (exp < -14) exp = -14;         // subnormal spacing 2^-24
  double ulp = std::ldexp(1.0, exp - 10);
  double r = std::nearbyint(a / ulp) * ulp;  // default rounding mode: nearest-even
  return std::copysign(r, x);
}

// bfloat16 round-to-nearest-even of a double (via the exact float value first when representable).
double round_bf16(double x) {
  if (!std::isfinite(x)) return x;
  double a = std::fabs(x);
  if (a == 0.0) return x;
  int e;
  std::frexp(a, &e);
  int exp = e - 1;
  if (exp < -126) exp = -126;
  double ulp = std::ldexp(1.0, exp - 7);
  double r = std::nearbyint(a / ulp) * ulp;
  return std::copysign(r, x);
}

// Inputs of the closure network: six rotation invariants of the strain rate S (tr S, the second and
// third principal invariants, tr S^2, tr S^3, |S| = sqrt(2 tr S^2); each as sign(x) log(1 + |x| /
// 1e-3^degree)), the wall distance and log(1 + |u| max(d, 1/2) / nu0). S: sxx, syy, szz, sxy, sxz, syz.
void mlp_features(const double S[6], double wd, double umag, double nu0, double x[8]) {
  const double sxx = S[0], syy = S[1], szz = S[2], sxy = S[3], sxz = S[4], syz = S[5];
  const double i1 = sxx + syy + szz;
  const double tr2 = sxx * sxx + syy * syy + szz * szz + 2.0 * (sxy * sxy + sxz * sxz + syz * syz);
  const double i2 = 0.5 * (i1 * i1 - tr2);
  const double i3 = sxx * (syy * szz - syz * syz) - sxy * (sxy * szz - syz * sxz) + sxz * (sxy * syz - syy * sxz);
  const double tr3 = i1 * i1 * i1 - 3.0 * i1 * i2 + 3.0 * i3;
  const double m = std::sqrt(2.0 * tr2);
  auto slog = [](double v, double ref) { return std::copysign(std::log1p(std::fabs(v) / ref), v); };
  const double r = 1e-3;
  x[0] = slog(i1, r);
  x[1] = slog(i2, r * r);
  x[2] = slog(i3, r * r * r);
  x[3] = slog(tr2, r * r);
  x[4] = slog(tr3, r * r * r);
  x[5] = slog(m, r);
  x[6] = wd;
  x[7] = std::log1p(umag * std::max(wd, 0.5) / nu0);
}

// Distance of v to the nearest BF16 rounding midpoint (where round-to-nearest changes its result).
double bf16_margin(double v) {
  const double a = std::fabs(v);
  if (!(a > 0.0) || !std::isfinite(a)) return INFINITY;
  int e;
  std::frexp(a, &e);                       // a in [2^(e-1), 2^e)
  const double spacing = std::ldexp(1.0, e - 8);  // 8 significant bits
  const double t = a / spacing;
  return std::fabs(t - std::floor(t) - 0.5) * spacing;
}

// The network on normalised inputs with the MFMA path's roundings: BF16 inputs, weights (all three
// layers) and hidden layers (after the bias, before the ReLU); accumulation and biases in FP64
// here. Returns nu_t = nu_scale * min(softplus(z), 10). margin (may be null): the smallest distance
// of a BF16-rounded input or positive hidden value to a rounding midpoint.
double mlp_forward(const or_mlp& net, const double xn[8], double* margin = nullptr) {
  const double* p = net.params;
  double h1[32], h2[32];
  double mg = INFINITY;
  if (margin)
    for (int i = 0; i < 8; i++) mg = std::min(mg, bf16_margin(xn[i]));
  for (int o = 0; o < 32; o++) {
    double a = p[256 + o];
    for (int i = 0; i < 8; i++) a += round_bf16(p[o * 8 + i]) * round_bf16(xn[i]);
    if (margin && a > 0.0) mg = std::min(mg, bf16_margin(a));
    h1[o] = std::max(round_bf16(a), 0.0);
  }
  for (int o = 0; o < 32; o++) {
    double a = p[1312 + o];
    for (int i = 0; i < 32; i++) a += round_bf16(p[288 + o * 32 + i]) * h1[i];
    if (margin && a > 0.0) mg = std::min(mg, bf16_margin(a));
    h2[o] = std::max(round_bf16(a), 0.0);
  }
  if (margin) *margin = mg;
  double z = p[1376];
  for (int i = 0; i < 32; i++) z += round_bf16(p[1344 + i]) * h2[i];
  const double sp = z > 20.0 ? z : std::log1p(std::exp(z));
  return net.nu_scale * std::min(sp, 10.0);
}

template <class T>
struct Solver {
  or_params prm;
  Lattice L;
  int nx, ny, nz;
  int64_t n;
  // derived block (SolverGPU.ts:160-187)
  T tau0, csmag2, uin, rhoIn, spongeStart, spongeNu, uGround, emaAlpha, uRef, tauWall, spongeIn, contactH;
  int ground, sideMode, collision, interp;
  int64_t refX;
  std::vector<uint32_t> flags, sdf;
  uint32_t wheel_count = 0;
  float wheels[32][4];
  std::vector<T> fA, fB;
  int parity = 0;
  bool pending_init = true;
  int64_t step = 0;
  std::vector<T> meanA, meanB, vel;
  // learned closure: network, wall distance (cells x 32) and the cells it runs on
  bool mlp_on = false;
  or_mlp net;
  std::vector<uint8_t> wdist, use_net;
  // per cell, the last step's BF16 rounding margin of the network evaluation (infinite off the network)
  mutable std::vector<double> margin;

  // Cells the closure network runs on (see or_mlp): fluid, and no cell of the pair group is solid
  // or has a plain pull (wgslLbm.ts:172) from a solid cell.
  void classify_net() {
    std::vector<uint8_t> bad(n, 0);
    const int Q = L.q, dim = L.dim;
    for (int64_t idx = 0; idx < n; idx++) {
      if (flags[idx] != 0u) { bad[idx] = 1; continue; }
      const int ix = (int)(idx % nx), iy = (int)((idx / nx) % ny), iz = (int)(idx / ((int64_t)nx * ny));
      for (int i = 1; i < Q; i++) {
        const int cx = L.c[i][0], cy = L.c[i][1], cz = L.c[i][2];
        int sx = ix - cx;
        const int sy = iy - cy, sz = dim == 3 ? iz - cz : 0;
        if (cx < 0) sx = std::min(sx, nx - 1);
        if (cx > 0 && sx < 0) continue;                                 // inlet
        if (sy < 0 || sy >= ny || sz < 0 || sz >= nz) continue;         // ground, mirrors, periodic
        if (flags[cell(sx, sy, sz)] != 0u) { bad[idx] = 1; break; }
      }
    }
    const int pair = (net.pair == 2 && nx % 2 == 0) ? 2 : 1;
    use_net.assign(n, 0);
    for (int64_t idx = 0; idx < n; idx++) {
      const int64_t g0 = idx - idx % pair;
      bool ok = true;
      for (int64_t j = g0; j < g0 + pair; j++) ok = ok && !bad[j];
      use_net[idx] = ok ? 1 : 0;
    }
  }

  void derive() {
    tau0 = T(3.0 * prm.nu + 0.5);
    csmag2 = T(prm.cs * prm.cs);
    uin = T(prm.u);
    rhoIn = T(1.0);
    ground = prm.ground;
    spongeStart = T(prm.sponge_start * nx);
    spongeNu = T(prm.sponge_nu);
    uGround = T(prm.ground == 2 ? prm.u : 0.0);
    emaAlpha = T(prm.ema_alpha);
    uRef = T(std::max(prm.u, 1e-4));
    refX = std::min<int64_t>(4, nx - 1);
    sideMode = prm.sides;
    collision = prm.collision;
    tauWall = T(prm.tau_wall);
    spongeIn = T(prm.sponge_in * nx);
    contactH = T(2.0);
    interp = prm.interp;
  }

  int64_t cell(int x, int y, int z) const { return (int64_t)x + (int64_t)nx * ((int64_t)y + (int64_t)ny * z); }
  const T* cur() const { return parity == 0 ? fA.data() : fB.data(); }
  T* cur_mut() { return parity == 0 ? fA.data() : fB.data(); }

  T store_round(T f, int i) const {
    switch (prm.storage) {
      case 1: return T((float)f);
      case 2: return T(L.w[i]) + T(round_half((double)f - L.w[i]));
      case 3: return T(L.w[i]) + T(round_bf16((double)f - L.w[i]));
      default: return f;
    }
  }

  // wgslLbm.ts:32-43
  T sdf_at(int64_t id) const {
    uint32_t s = sdf[id];
    if (s == 0u) return T(-1);
    uint32_t b = ~s;
    float f;
    std::memcpy(&f, &b, 4);
    return T(f);
  }
  T link_q(int64_t fluid, int64_t solid) const {
    if (interp == 0) return T(0.5);
    T a = sdf_at(fluid), b = sdf_at(solid);
    if (a < 0 || b < 0 || a + b < T(1e-6)) return T(0.5);
    return std::min(std::max(a / (a + b), T(0.01)), T(0.99));
  }
  // wgslLbm.ts:21-28
  void wheel_vel(uint32_t k, const T p[3], T out[3]) const {
    const float* a = wheels[2 * k];
    const float* b = wheels[2 * k + 1];
    T d[3] = {p[0] - T(a[0]), p[1] - T(a[1]), p[2] - T(a[2])};
    T cr[3] = {T(b[1]) * d[2] - T(b[2]) * d[1], T(b[2]) * d[0] - T(b[0]) * d[2], T(b[0]) * d[1] - T(b[1]) * d[0]};
    for (int j = 0; j < 3; j++) out[j] = uGround * T(a[3]) * cr[j];
  }

  // init kernel, wgslLbm.ts:296-341: equilibrium at (1, (U,0,0)) on fluid cells, at rest on solids.
  void do_init() {
    for (int b = 0; b < 2; b++) {
      T* f = b == 0 ? fA.data() : fB.data();
      for (int64_t idx = 0; idx < n; idx++) {
        T U = flags[idx] != 0u ? T(0) : uin;
        for (int i = 0; i < L.q; i++) {
          T cx = T(L.c[i][0]);
          T v = T(L.w[i]) * (T(1) + T(3) * cx * U + T(4.5) * cx * cx * U * U - T(1.5) * U * U);
          f[(int64_t)i * n + idx] = store_round(v, i);
        }
      }
    }
    pending_init = false;
  }

  // One cell of streamCollideBody (wgslLbm.ts:134-293). Writes the post-collision populations.
  void update_cell(int64_t idx, const T* fin, T* out, bool periodic_x) const {
    const int Q = L.q, dim = L.dim;
    const int NX = nx, NY = ny, NZ = nz;
    int ix = (int)(idx % nx);
    int iy = (int)((idx / nx) % ny);
    int iz = (int)(idx / ((int64_t)nx * ny));
    const T U = uin;
    bool wall = false, gwall = false;
    T mvSum = 0, mvW = 0;
    uint32_t wm = 0;
    // tyre contact patch (wgslLbm.ts:276-287)
    T belt = ground == 2 ? T(1) : T(0);
    if (belt > 0 && iy == 0) {
      int r = (int)contactH;
      for (int dz = -r; dz <= r; dz++)
        for (int dx = -r; dx <= r; dx++)
          for (int dy = 0; dy <= 1; dy++) {
            int qx = ix + dx, qz = iz + dz;
            if (qx >= 0 && qx < NX && qz >= 0 && qz < NZ) {
              uint32_t fq = flags[cell(qx, dy, qz)];
              if (fq != 0u && (fq >> 8) == 0u) belt = 0;
            }
          }
    }
    if (flags[idx] != 0u) {  // solid: rest populations (wgslLbm.ts:95-97)
      for (int i = 0; i < Q; i++) out[i] = T(L.w[i]);
      return;
    }
    T f[19];
    f[0] = fin[idx];
    for (int i = 1; i < Q; i++) {
      const int cx = L.c[i][0], cy = L.c[i][1], cz = L.c[i][2];
      const T w = T(L.w[i]);
      int sx = ix - cx, sy = iy - cy, sz = dim == 3 ? iz - cz : 0;
      if (periodic_x) {
        sx = (sx + NX) % NX;
      } else if (cx < 0) {
        sx = std::min(sx, NX - 1);
      }
      bool yOut = cy > 0 ? sy < 0 : (cy < 0 ? sy >= NY : false);
      bool zOut = dim == 3 ? (cz > 0 ? sz < 0 : (cz < 0 ? sz >= NZ : false)) : false;
      bool inl = !periodic_x && cx > 0 && sx < 0;
      T v;
      if (inl) {
        T cu = T(cx) * U;
        v = w * rhoIn * (T(1) + T(3) * cu + T(4.5) * cu * cu - T(1.5) * U * U);
      } else if (cy > 0 && yOut && ground != 0) {
        wall = true; gwall = true;
        v = fin[(int64_t)L.opp[i] * n + idx];
        if (cx != 0) v += belt * (T(6.0 * L.w[i] * cx) * uGround);
      } else {
        int sxc = std::min(std::max(sx, 0), NX - 1);
        if (dim == 3 && yOut && zOut) {
          v = fin[(int64_t)L.mYZ[i] * n + cell(sxc, iy, iz)];
        } else if (yOut) {
          v = fin[(int64_t)L.mY[i] * n + cell(sxc, iy, dim == 3 ? std::min(std::max(sz, 0), NZ - 1) : 0)];
        } else if (dim == 3 && zOut) {
          if (sideMode == 1) v = fin[(int64_t)i * n + cell(sxc, sy, (sz + NZ) % NZ)];
          else v = fin[(int64_t)L.mZ[i] * n + cell(sxc, sy, iz)];
        } else {
          int64_t sc = cell(sxc, sy, sz);
          uint32_t fs = flags[sc];
          if (fs != 0u) {
            wall = true;
            T fb = fin[(int64_t)L.opp[i] * n + idx];
            T q = link_q(idx, sc);
            v = fb;
            if (q < T(0.5)) {
              int ax = ix + cx, ay = iy + cy, az = dim == 3 ? iz + cz : 0;
              if (ax >= 0 && ay >= 0 && az >= 0 && ax < NX && ay < NY && az < NZ && flags[cell(ax, ay, az)] == 0u)
                v = T(2) * q * fb + (T(1) - T(2) * q) * fin[(int64_t)L.opp[i] * n + cell(ax, ay, az)];
              else
                q = T(0.5);
            } else if (q > T(0.5)) {
              v = (T(0.5) / q) * fb + (T(1) - T(0.5) / q) * fin[(int64_t)i * n + idx];
            }
            if ((fs >> 8) != 0u) {
              T p[3] = {T(0.5) * T(ix + sxc) + T(0.5), T(0.5) * T(iy + sy) + T(0.5), T(0.5) * T(iz + sz) + T(0.5)};
              T uw[3];
              wheel_vel((fs >> 8) - 1u, p, uw);
              T mv = T(6.0 * L.w[i]) * (T(cx) * uw[0] + T(cy) * uw[1] + T(cz) * uw[2]) * (q > T(0.5) ? T(0.5) / q : T(1));
              v += mv; mvSum += mv; mvW += w; wm |= 1u << i;
            }
          } else {
            v = fin[(int64_t)i * n + sc];
          }
        }
      }
      f[i] = v;
    }
    if (mvW > 0) {  // wgslLbm.ts:289-291
      T corr = mvSum / mvW;
      for (int i = 1; i < Q; i++)
        if (wm & (1u << i)) f[i] -= T(L.w[i]) * corr;
    }
    // collision (wgslLbm.ts:177-240)
    T rho = 0, j[3] = {0, 0, 0}, S[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int i = 0; i < Q; i++) {
      rho += f[i];
      for (int a = 0; a < 3; a++) {
        j[a] += T(L.c[i][a]) * f[i];
        for (int b = 0; b < 3; b++) S[a][b] += T(L.c[i][a] * L.c[i][b]) * f[i];
      }
    }
    T irho = T(1) / rho;
    T u[3] = {j[0] * irho, j[1] * irho, dim == 3 ? j[2] * irho : T(0)};
    T usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    T Pi[3][3];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) Pi[a][b] = S[a][b] - rho * (u[a] * u[b] + (a == b ? T(1.0 / 3.0) : T(0)));
    T pn = 0, tr = 0;
    for (int a = 0; a < dim; a++) {
      tr += Pi[a][a];
      for (int b = 0; b < dim; b++) pn += Pi[a][b] * Pi[a][b];
    }
    T piN = std::sqrt(pn);
    T fx = T(ix);
    T sp = std::min(std::max((fx - spongeStart) / std::max(T(NX) - spongeStart, T(1)), T(0)), T(1));
    if (periodic_x) sp = 0;
    T t0 = tau0 + T(3) * spongeNu * sp * sp;
    if (gwall) t0 = std::max(t0, tauWall);
    T tau = T(0.5) * (t0 + std::sqrt(t0 * t0 + T(18.0 * std::sqrt(2.0)) * csmag2 * piN * irho));
    if (mlp_on && dim == 3 && use_net[idx]) {
      // learned closure: strain rate from the non-equilibrium stress with the Smagorinsky tau,
      // S = -3 Pi^neq / (2 rho tau); nu_t = max(network, floor x nu_smag); a non-finite (or huge)
      // output keeps Smagorinsky
      const double sc = -1.5 / ((double)rho * (double)tau);
      const double S6[6] = {sc * (double)Pi[0][0], sc * (double)Pi[1][1], sc * (double)Pi[2][2],
                            sc * (double)Pi[0][1], sc * (double)Pi[0][2], sc * (double)Pi[1][2]};
      double x[8], xn[8];
      mlp_features(S6, wdist[idx] / 32.0, std::sqrt((double)usq), prm.nu, x);
      for (int k = 0; k < 8; k++) xn[k] = (x[k] - net.mu[k]) / (net.sigma[k] > 0 ? net.sigma[k] : 1.0);
      const double nu = mlp_forward(net, xn, &margin[idx]);
      if (nu == nu && nu < 1e3) {
        const double nus = ((double)tau - (double)t0) / 3.0;
        tau = t0 + T(3.0 * std::max(nu, net.smag_floor * nus));
      }
    }
    T k1 = (T(1) - T(1) / tau) * T(4.5);
    bool useBGK = wall || collision == 1;
    T si = std::min(std::max(T(1) - fx / std::max(spongeIn, T(1)), T(0)), T(1));
    T sIn = spongeIn > 0 ? T(0.08) * si * si : T(0);
    T trP = tr * T(1.0 / 3.0);
    for (int i = 0; i < Q; i++) {
      const T w = T(L.w[i]);
      T ci[3] = {T(L.c[i][0]), T(L.c[i][1]), T(L.c[i][2])};
      T cu = ci[0] * u[0] + ci[1] * u[1] + ci[2] * u[2];
      T feq = w * rho * (T(1) + T(3) * cu + T(4.5) * cu * cu - T(1.5) * usq);
      T qp = 0;
      for (int a = 0; a < dim; a++)
        for (int b = 0; b < dim; b++) qp += ci[a] * ci[b] * Pi[a][b];
      T fr = w * k1 * (qp - trP);
      T fo = feq + (useBGK ? (T(1) - T(1) / tau) * (f[i] - feq) : fr);
      if (sIn > 0) {
        T tgt = w * (T(1) + T(3) * ci[0] * U + T(4.5) * ci[0] * ci[0] * U * U - T(1.5) * U * U);
        fo = fo * (T(1) - sIn) + tgt * sIn;
      }
      out[i] = fo;
    }
  }

  void do_step(int steps) {
    if (pending_init) do_init();
    if (mlp_on) classify_net();
    margin.assign(n, INFINITY);
    const int Q = L.q;
    for (int s = 0; s < steps; s++) {
      const T* fin = cur();
      T* fout = parity == 0 ? fB.data() : fA.data();
      const bool px = prm.periodic_x != 0;
#pragma omp parallel for schedule(static) num_threads(prm.threads > 0 ? prm.threads : omp_threads())
      for (int64_t idx = 0; idx < n; idx++) {
        T o[19];
        update_cell(idx, fin, o, px);
        for (int i = 0; i < Q; i++) fout[(int64_t)i * n + idx] = store_round(o[i], i);
      }
      parity ^= 1;
      step++;
    }
  }

  static int omp_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
  }

  // Macro pass + final reduction (wgslLbm.ts:353-521). out[0..3] = (Fx,Fy,Fz,max u²),
  // out[4..7] = (Σρ_ref, n_ref, nan count, solid count).
  void sample(double out[8]) {
    if (pending_init) do_init();
    const int Q = L.q, dim = L.dim;
    const T* fin = cur();
    const int NX = nx, NY = ny, NZ = nz;
    auto safe = [&](int x, int y, int z) -> int64_t {
      if (x < 0 || y < 0 || z < 0 || x >= NX || y >= NY || z >= NZ) return -1;
      return cell(x, y, z);
    };
    double A0 = 0, A1 = 0, A2 = 0, Bs[4] = {0, 0, 0, 0};
    double Amax = 0;
    for (int64_t idx = 0; idx < n; idx++) {
      int ix = (int)(idx % nx), iy = (int)((idx / nx) % ny), iz = (int)(idx / ((int64_t)nx * ny));
      T outv[4] = {0, 0, 0, 1};
      T A[4] = {0, 0, 0, 0}, B[4] = {0, 0, 0, 0};
      if (flags[idx] == 0u) {
        T f[19];
        for (int i = 0; i < Q; i++) f[i] = fin[(int64_t)i * n + idx];
        T rho = 0, j[3] = {0, 0, 0};
        for (int i = 0; i < Q; i++) {
          rho += f[i];
          for (int a = 0; a < dim; a++) j[a] += T(L.c[i][a]) * f[i];
        }
        T u[3] = {j[0] / rho, j[1] / rho, dim == 3 ? j[2] / rho : T(0)};
        outv[0] = u[0]; outv[1] = u[1]; outv[2] = u[2]; outv[3] = rho;
        T F[3] = {0, 0, 0}, mvSum = 0, mvW = 0, Cw[3] = {0, 0, 0};
        for (int i = 1; i < Q; i++) {
          const int cx = L.c[i][0], cy = L.c[i][1], cz = L.c[i][2];
          int64_t s = safe(ix + cx, iy + cy, iz + cz);
          if (s < 0 || flags[s] == 0u) continue;
          T q = link_q(idx, s);
          T fr = f[i];
          if (q < T(0.5)) {
            int64_t bk = safe(ix - cx, iy - cy, iz - cz);
            if (bk >= 0 && flags[bk] == 0u) fr = T(2) * q * f[i] + (T(1) - T(2) * q) * fin[(int64_t)i * n + bk];
            else q = T(0.5);
          } else if (q > T(0.5)) {
            fr = (T(0.5) / q) * f[i] + (T(1) - T(0.5) / q) * f[L.opp[i]];
          }
          uint32_t tg = flags[s] >> 8;
          if (tg != 0u) {
            T p[3] = {T(ix) + T(0.5 + cx / 2.0), T(iy) + T(0.5 + cy / 2.0), T(iz) + T(0.5 + cz / 2.0)};
            T uw[3];
            wheel_vel(tg - 1u, p, uw);
            T mv = -T(6.0 * L.w[i]) * (q > T(0.5) ? T(0.5) / q : T(1)) * (T(cx) * uw[0] + T(cy) * uw[1] + T(cz) * uw[2]);
            fr += mv; mvSum += mv; mvW += T(L.w[i]);
            Cw[0] += T(cx) * T(L.w[i]); Cw[1] += T(cy) * T(L.w[i]); Cw[2] += T(cz) * T(L.w[i]);
          }
          T m = f[i] + fr - T(2.0 * L.w[i]);
          F[0] += T(cx) * m; F[1] += T(cy) * m; F[2] += T(cz) * m;
        }
        if (mvW > 0) {
          T r = mvSum / mvW;
          for (int a = 0; a < 3; a++) F[a] -= Cw[a] * r;
        }
        T uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        bool bad = !(uu < T(1e6)) || !(rho > 0);
        A[0] = F[0]; A[1] = F[1]; A[2] = F[2]; A[3] = bad ? T(1e6) : uu;
        B[0] = ix == refX ? rho : T(0);
        B[1] = ix == refX ? T(1) : T(0);
        B[2] = bad ? T(1) : T(0);
        for (int k = 0; k < 4; k++) meanA[4 * idx + k] = meanA[4 * idx + k] * (T(1) - emaAlpha) + outv[k] * emaAlpha;
        meanB[idx] = meanB[idx] * (T(1) - emaAlpha) + uu * emaAlpha;
      } else {
        T rs = 0, rn = 0;
        const int nb[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
        for (int k = 0; k < (dim == 3 ? 6 : 4); k++) {
          int64_t s = safe(ix + nb[k][0], iy + nb[k][1], iz + nb[k][2]);
          if (s >= 0 && flags[s] == 0u) {
            T r = 0;
            for (int i = 0; i < Q; i++) r += fin[(int64_t)i * n + s];
            rs += r; rn += 1;
          }
        }
        if (rn > 0) outv[3] = rs / rn;
        for (int k = 0; k < 4; k++) meanA[4 * idx + k] = meanA[4 * idx + k] * (T(1) - emaAlpha) + outv[k] * emaAlpha;
        meanB[idx] = 0;
        B[3] = 1;
      }
      for (int k = 0; k < 4; k++) vel[4 * idx + k] = outv[k];
      A0 += (double)A[0]; A1 += (double)A[1]; A2 += (double)A[2];
      Amax = std::max(Amax, (double)A[3]);
      for (int k = 0; k < 4; k++) Bs[k] += (double)B[k];
    }
    out[0] = A0; out[1] = A1; out[2] = A2; out[3] = Amax;
    for (int k = 0; k < 4; k++) out[4 + k] = Bs[k];
  }

  void clear_stats() {
    for (int64_t i = 0; i < n; i++) {
      meanA[4 * i] = meanA[4 * i + 1] = meanA[4 * i + 2] = 0;
      meanA[4 * i + 3] = 1;
      meanB[i] = 0;
    }
  }
};

struct Handle {
  bool dbl;
  Solver<double>* d = nullptr;
  Solver<float>* f = nullptr;
};

template <class T>
Solver<T>* make_solver(const or_params* p) {
  auto* s = new Solver<T>();
  s->prm = *p;
  s->nx = p->nx; s->ny = p->ny; s->nz = p->nz;
  s->n = (int64_t)p->nx * p->ny * p->nz;
  s->L = make_lattice(p->nz > 1);
  s->derive();
  s->flags.assign(s->n, 0u);
  s->sdf.assign(s->n, 0u);
  std::memset(s->wheels, 0, sizeof(s->wheels));
  s->fA.assign((size_t)s->L.q * s->n, T(0));
  s->fB.assign((size_t)s->L.q * s->n, T(0));
  s->meanA.assign(4 * s->n, T(0));
  s->meanB.assign(s->n, T(0));
  s->vel.assign(4 * s->n, T(0));
  s->clear_stats();
  s->pending_init = true;
  return s;
}

}  // namespace

extern "C" {

// precision_t: 0 = FP64 arithmetic (the oracle), 1 = FP32 arithmetic (CPU-baseline timing).
void* or_create(const or_params* p, int fp32) {
  auto* h = new Handle();
  h->dbl = !fp32;
  if (h->dbl) h->d = make_solver<double>(p);
  else h->f = make_solver<float>(p);
  return h;
}

void or_destroy(void* hp) {
  auto* h = (Handle*)hp;
  delete h->d;
  delete h->f;
  delete h;
}

#define OR_DISPATCH(...) \
  do { auto* h = (Handle*)hp; if (h->dbl) { auto* s = h->d; __VA_ARGS__; } else { auto* s = h->f; __VA_ARGS__; } } while (0)

int or_q(void* hp) { int q = 0; OR_DISPATCH(q = s->L.q); return q; }

void or_set_params(void* hp, const or_params* p) {
  OR_DISPATCH({
    int st = s->prm.storage, px = s->prm.periodic_x;
    s->prm = *p;
    s->prm.storage = st;
    s->prm.periodic_x = px;
    s->prm.nx = s->nx; s->prm.ny = s->ny; s->prm.nz = s->nz;
    s->derive();
  });
}

void or_upload_flags(void* hp, const uint32_t* f) { OR_DISPATCH(std::memcpy(s->flags.data(), f, 4 * s->n)); }

// SolverGPU.ts:312-322: values < 0 (or NaN) mean unknown (stored as 0), else ~bits(f32(d)).
void or_upload_sdf(void* hp, const float* d) {
  OR_DISPATCH({
    for (int64_t i = 0; i < s->n; i++) {
      if (!(d[i] >= 0.0f)) { s->sdf[i] = 0u; continue; }
      uint32_t b;
      std::memcpy(&b, &d[i], 4);
      s->sdf[i] = ~b;
    }
  });
}

// wheels: 8 floats per wheel (cx, cy, cz, axis x, y, z, r, hw); SolverGPU.ts:325-335.
void or_set_wheels(void* hp, const float* w, int k) {
  OR_DISPATCH({
    std::memset(s->wheels, 0, sizeof(s->wheels));
    int m = std::min(std::max(k, 0), 16);
    s->wheel_count = (uint32_t)m;
    for (int j = 0; j < m; j++) {
      const float* e = w + 8 * j;
      s->wheels[2 * j][0] = e[0]; s->wheels[2 * j][1] = e[1]; s->wheels[2 * j][2] = e[2];
      s->wheels[2 * j][3] = 1.0f / std::max(e[6], 1e-3f);
      s->wheels[2 * j + 1][0] = e[3]; s->wheels[2 * j + 1][1] = e[4]; s->wheels[2 * j + 1][2] = e[5];
      s->wheels[2 * j + 1][3] = e[7];
    }
  });
}

void or_reset(void* hp) { OR_DISPATCH({ s->pending_init = true; s->step = 0; s->clear_stats(); }); }
void or_clear_stats(void* hp) { OR_DISPATCH(s->clear_stats()); }
void or_step(void* hp, int steps) { OR_DISPATCH(s->do_step(steps)); }
void or_sample(void* hp, double* out) { OR_DISPATCH(s->sample(out)); }
int64_t or_step_count(void* hp) { int64_t v = 0; OR_DISPATCH(v = s->step); return v; }

// Current populations as full values f (not deviations), Q*N, direction-major.
void or_read_populations(void* hp, double* out) {
  OR_DISPATCH({
    if (s->pending_init) s->do_init();
    const auto* f = s->cur();
    for (size_t i = 0; i < (size_t)s->L.q * s->n; i++) out[i] = (double)f[i];
  });
}

// Overwrite the current populations (after any pending init), rounded to the storage format.
void or_upload_populations(void* hp, const double* in) {
  OR_DISPATCH({
    if (s->pending_init) s->do_init();
    auto* f = s->cur_mut();
    for (int i = 0; i < s->L.q; i++)
      for (int64_t c = 0; c < s->n; c++) f[(int64_t)i * s->n + c] = s->store_round((decltype(s->tau0))in[(int64_t)i * s->n + c], i);
  });
}

// Macro fields of the last sample: vel (ux,uy,uz,rho) and EMA meanA (4/cell) and meanB (1/cell).
void or_read_macro(void* hp, double* vel, double* meanA, double* meanB) {
  OR_DISPATCH({
    for (int64_t i = 0; i < 4 * s->n; i++) { vel[i] = (double)s->vel[i]; meanA[i] = (double)s->meanA[i]; }
    for (int64_t i = 0; i < s->n; i++) meanB[i] = (double)s->meanB[i];
  });
}

// Attach the learned closure (net == NULL detaches it). wdist: n bytes, wall distance x 32.
void or_set_mlp(void* hp, const or_mlp* net, const uint8_t* wdist) {
  OR_DISPATCH({
    if (!net) {
      s->mlp_on = false;
    } else {
      s->net = *net;
      s->wdist.assign(wdist, wdist + s->n);
      s->mlp_on = true;
    }
  });
}

// Per cell, the BF16 rounding margin of the last step's network evaluation (infinite where the
// network did not run): where it is below the FP32 error of the device's inputs, the device may
// round that input the other way.
void or_mlp_margins(void* hp, double* out) {
  OR_DISPATCH({
    for (int64_t i = 0; i < s->n; i++) out[i] = (size_t)i < s->margin.size() ? s->margin[i] : INFINITY;
  });
}

// The closure network alone (n x 8 raw inputs -> n eddy viscosities), for checking the restatement.
void or_mlp_infer(const or_mlp* net, const double* raw, int64_t n, double* nu) {
  for (int64_t c = 0; c < n; c++) {
    double xn[8];
    for (int k = 0; k < 8; k++) xn[k] = (raw[8 * c + k] - net->mu[k]) / (net->sigma[k] > 0 ? net->sigma[k] : 1.0);
    nu[c] = mlp_forward(*net, xn);
  }
}

// mlp_features for n cells (S: n x 6), for checking the restatement.
void or_mlp_features(const double* S, const double* wd, const double* umag, int64_t n, double nu0, double* x) {
  for (int64_t c = 0; c < n; c++) mlp_features(S + 6 * c, wd[c], umag[c], nu0, x + 8 * c);
}

double or_round_half(double x) { return round_half(x); }
double or_round_bf16(double x) { return round_bf16(x); }

}  // extern "C"
