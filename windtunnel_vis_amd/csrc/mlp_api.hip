// C-ABI entry points of the learned subgrid closure: loading a network into a solver, wall
// distances, training datasets generated from fine runs, GPU training and stand-alone inference.
// Declared (with their contracts) in include/wt.h; no reference counterpart.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <random>
#include <vector>

#include "solver.hpp"

// Device scratch of wt_mlp_infer_noalloc, which calls no allocator: the network in the MFMA layout,
// the staged raw inputs and the results of at most WT_MLP_NOALLOC_MAX cells. One copy per device
// (a __device__ array exists once on every GPU the library runs on, for the life of the process).
namespace {
constexpr size_t NOALLOC_NET_BYTES = (sizeof(wt::MlpDev) + 255) / 256 * 256;
constexpr size_t NOALLOC_RAW_BYTES = sizeof(float) * wt::MLP_IN * (size_t)WT_MLP_NOALLOC_MAX;
constexpr size_t NOALLOC_NU_BYTES = sizeof(float) * (size_t)WT_MLP_NOALLOC_MAX;
constexpr size_t NOALLOC_BYTES = NOALLOC_NET_BYTES + NOALLOC_RAW_BYTES + NOALLOC_NU_BYTES;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

// The scratch array's device address, per device, looked up once with that device current. The
// array's name is only a host-side stand-in for the symbol: host code uses this address.
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

int scratch_for_current_device(unsigned char** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return wt::fail(WT_E_NODEV, "device index out of range for the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return wt::fail(WT_E_HIP, "hipGetSymbolAddress returned a null scratch address");
    g_scratch_ptr[dev] = (unsigned char*)p;
  }
  *out = g_scratch_ptr[dev];
  return WT_OK;
}
}  // namespace

using namespace wt;

namespace {

short host_bf16(float x) {  // round to nearest even, as to_bf16 on the device
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// FP32 master parameters + normalisation -> the MFMA layout that mlp_wave_z reads.
void build_dev(const wt_mlp* net, MlpDev* d) {
  std::memset(d, 0, sizeof(*d));
  const float* p = net->params;
  for (int l = 0; l < 64; l++) {
    const int row = l % 32, h = l / 32;
    for (int t = 0; t < 4; t++) d->lane[l].w1a[t] = host_bf16(p[MLP_OFF_W1 + row * MLP_IN + 4 * h + t]);
    for (int c = 0; c < 2; c++)
      for (int j = 0; j < 8; j++) d->lane[l].w2b[c][j] = host_bf16(p[MLP_OFF_W2 + row * MLP_H + mlp_hidden_of(c, h, j)]);
  }
  for (int h = 0; h < 2; h++)
    for (int r = 0; r < 16; r++) {
      const int i = 8 * (r / 4) + 4 * h + (r % 4);
      d->half[h].b1d[r] = p[MLP_OFF_B1 + i];
      d->half[h].b2d[r] = p[MLP_OFF_B2 + i];
    }
  for (int c = 0; c < 2; c++)
    for (int h = 0; h < 2; h++)
      for (int j = 0; j < 8; j++) d->half[h].w3b[c][j] = host_bf16(p[MLP_OFF_W3 + mlp_hidden_of(c, h, j)]);
  d->b3 = p[MLP_OFF_B3];
  for (int i = 0; i < MLP_IN; i++) {
    const double isig = net->sigma[i] > 0.0f ? 1.0 / net->sigma[i] : 1.0;
    d->norm.mu[i] = net->mu[i];
    d->norm.isig[i] = (float)isig;
    // the step kernel's raw inputs: base-2 logarithms, except the wall distance (its byte, 32 per cell)
    const double unit = i == 6 ? 1.0 / 32.0 : 0.69314718055994530942;
    d->norm.na[i] = (float)(unit * isig);
    d->norm.nb[i] = (float)(-(double)net->mu[i] * isig);
  }
  d->norm.nu_scale = net->nu_scale;
  d->norm.smag_floor = net->smag_floor;
}

MlpNorm host_norm(const wt_mlp* net) {
  MlpNorm m;
  std::memset(&m, 0, sizeof(m));
  for (int i = 0; i < MLP_IN; i++) {
    m.mu[i] = net->mu[i];
    m.isig[i] = net->sigma[i] > 0.0f ? 1.0f / net->sigma[i] : 1.0f;
  }
  m.nu_scale = net->nu_scale > 0.0f ? net->nu_scale : 1.0f;
  m.smag_floor = net->smag_floor;
  return m;
}

}  // namespace

// Wall distance of a solver's cells (bytes, 32 per cell, capped at 4 cells); built when missing.
int wt::ensure_wdist(wt_solver* s) {
  if (s->wdist && s->wdist_valid) return WT_OK;
  if (!s->wdist) HIP_TRY(hipMalloc((void**)&s->wdist, (size_t)s->n + 64));
  uint32_t* tmp = nullptr;
  HIP_TRY(hipMalloc((void**)&tmp, sizeof(uint32_t) * (size_t)s->n));
  launch_wall_distance(s->stream, s->flags, tmp, s->wdist, s->dims.nx, s->dims.ny, s->dims.nz);
  const hipError_t e = hipStreamSynchronize(s->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(WT_E_HIP, std::string("wall distance: ") + hipGetErrorString(e));
  s->wdist_valid = true;
  return WT_OK;
}

struct wt_dataset {
  float* data = nullptr;  // records of 10 floats
  int64_t cap = 0;
  int64_t count = 0;
};

extern "C" {

int wt_mlp_load(wt_solver* s, const wt_mlp* net) {
  if (!s || !net) return fail(WT_E_ARG, "null argument");
  if (s->q != 19) return fail(WT_E_ARG, "the learned closure is three-dimensional (D3Q19)");
  MlpDev h;
  build_dev(net, &h);
  HIP_TRY(hipStreamSynchronize(s->stream));  // steps in flight may read the previous network
  if (!s->mlp) HIP_TRY(hipMalloc((void**)&s->mlp, sizeof(MlpDev)));
  HIP_TRY(hipMemcpy(s->mlp, &h, sizeof(MlpDev), hipMemcpyHostToDevice));
  return ensure_wdist(s);
}

int wt_read_wall_distance(wt_solver* s, uint8_t* out) {
  if (!s || !out) return fail(WT_E_ARG, "null argument");
  if (int r = ensure_wdist(s)) return r;
  HIP_TRY(hipMemcpy(out, s->wdist, (size_t)s->n, hipMemcpyDeviceToHost));
  return WT_OK;
}

int wt_mlp_infer(const wt_mlp* net, const float* raw, int64_t n, float* nu) {
  if (!net || (n > 0 && (!raw || !nu)) || n < 0) return fail(WT_E_ARG, "bad inference arguments");
  if (n == 0) return WT_OK;
  MlpDev h;
  build_dev(net, &h);
  MlpDev* d = nullptr;
  float *draw = nullptr, *dnu = nullptr;
  hipStream_t st = nullptr;
  int rc = WT_OK;
  if (hipMalloc((void**)&d, sizeof(MlpDev)) != hipSuccess || hipMalloc((void**)&draw, sizeof(float) * MLP_IN * (size_t)n) != hipSuccess ||
      hipMalloc((void**)&dnu, sizeof(float) * (size_t)n) != hipSuccess || hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
    rc = fail(WT_E_NOMEM, "inference buffers");
  } else {
    hipError_t e = hipMemcpyAsync(d, &h, sizeof(MlpDev), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(draw, raw, sizeof(float) * MLP_IN * (size_t)n, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
      launch_mlp_infer(st, d, draw, n, dnu);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(nu, dnu, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = fail(WT_E_HIP, std::string("inference: ") + hipGetErrorString(e));
  }
  if (st) (void)hipStreamDestroy(st);
  if (d) (void)hipFree(d);
  if (draw) (void)hipFree(draw);
  if (dnu) (void)hipFree(dnu);
  return rc;
}

int wt_mlp_infer_noalloc(const wt_mlp* net, const float* raw, int64_t n, float* nu) {
  if (!net || (n > 0 && (!raw || !nu)) || n < 0) return fail(WT_E_ARG, "bad inference arguments");
  if (n > WT_MLP_NOALLOC_MAX)
    return fail(WT_E_ARG, "wt_mlp_infer_noalloc: at most WT_MLP_NOALLOC_MAX (65536) cells per call; split the input or use wt_mlp_infer");
  if (n == 0) return WT_OK;
  unsigned char* base = nullptr;
  if (int r = scratch_for_current_device(&base)) return r;
  MlpDev h;
  build_dev(net, &h);
  MlpDev* dnet = (MlpDev*)base;
  float* draw = (float*)(base + NOALLOC_NET_BYTES);
  float* dnu = (float*)(base + NOALLOC_NET_BYTES + NOALLOC_RAW_BYTES);
  // the null stream, synchronised before returning: the staged inputs and results are in the
  // shared scratch, so the next call (or another thread's) may overwrite them afterwards
  hipStream_t st = nullptr;
  hipError_t e = hipMemcpyAsync(dnet, &h, sizeof(MlpDev), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(draw, raw, sizeof(float) * MLP_IN * (size_t)n, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    launch_mlp_infer(st, dnet, draw, n, dnu);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(nu, dnu, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail(WT_E_HIP, std::string("inference: ") + hipGetErrorString(e));
  return WT_OK;
}

int wt_dataset_create(int64_t capacity, wt_dataset** out) {
  if (!out || capacity < 1) return fail(WT_E_ARG, "dataset capacity must be >= 1");
  auto* d = new wt_dataset();
  if (hipMalloc((void**)&d->data, sizeof(float) * 10 * (size_t)capacity) != hipSuccess) {
    delete d;
    return fail(WT_E_NOMEM, "dataset allocation failed");
  }
  d->cap = capacity;
  *out = d;
  return WT_OK;
}

int wt_dataset_destroy(wt_dataset* d) {
  if (!d) return WT_OK;
  if (d->data) (void)hipFree(d->data);
  delete d;
  return WT_OK;
}

int wt_dataset_count(const wt_dataset* d, int64_t* n) {
  if (!d || !n) return fail(WT_E_ARG, "null argument");
  *n = d->count;
  return WT_OK;
}

int wt_dataset_read(const wt_dataset* d, int64_t first, int64_t count, float* out) {
  if (!d || (count > 0 && !out) || first < 0 || count < 0 || first + count > d->count) return fail(WT_E_ARG, "bad dataset range");
  if (count == 0) return WT_OK;
  HIP_TRY(hipMemcpy(out, d->data + first * 10, sizeof(float) * 10 * (size_t)count, hipMemcpyDeviceToHost));
  return WT_OK;
}

int wt_dataset_append(wt_dataset* d, const float* recs, int64_t count) {
  if (!d || (count > 0 && !recs) || count < 0) return fail(WT_E_ARG, "bad dataset records");
  if (d->count + count > d->cap) return fail(WT_E_NOMEM, "dataset capacity exceeded");
  if (count == 0) return WT_OK;
  HIP_TRY(hipMemcpy(d->data + d->count * 10, recs, sizeof(float) * 10 * (size_t)count, hipMemcpyHostToDevice));
  d->count += count;
  return WT_OK;
}

int wt_dataset_add(wt_dataset* d, wt_solver* fine, wt_solver* coarse, int64_t* added) {
  if (!d || !fine || !coarse) return fail(WT_E_ARG, "null argument");
  if (fine->q != 19 || coarse->q != 19) return fail(WT_E_ARG, "datasets come from D3Q19 runs");
  if (fine->dims.nx != 2 * coarse->dims.nx || fine->dims.ny != 2 * coarse->dims.ny || fine->dims.nz != 2 * coarse->dims.nz)
    return fail(WT_E_ARG, "the fine grid must be exactly twice the coarse grid in every direction");
  if (fine->pending_init) {
    if (int r = run_init(fine)) return r;
  }
  if (int r = ensure_wdist(coarse)) return r;
  HIP_TRY(hipStreamSynchronize(fine->stream));
  const int cnx = coarse->dims.nx, cny = coarse->dims.ny, cnz = coarse->dims.nz;
  const int64_t nc = coarse->n;
  const int64_t nb = coarse_blocks(nc);
  const float nu0 = (float)coarse->prm.nu;
  float* cf = nullptr;
  float2* nd = nullptr;
  uint8_t* bad = nullptr;
  uint32_t* counts = nullptr;
  uint64_t* offs = nullptr;
  int rc = WT_OK;
  hipStream_t st = fine->stream;
  if (hipMalloc((void**)&cf, sizeof(float) * 9 * (size_t)nc) != hipSuccess || hipMalloc((void**)&bad, (size_t)nc) != hipSuccess ||
      hipMalloc((void**)&nd, sizeof(float2) * (size_t)nc) != hipSuccess ||
      hipMalloc((void**)&counts, sizeof(uint32_t) * (size_t)nb) != hipSuccess ||
      hipMalloc((void**)&offs, sizeof(uint64_t) * (size_t)nb) != hipSuccess) {
    rc = fail(WT_E_NOMEM, "dataset temporaries");
  } else {
    launch_coarse_fields(st, fine->prec, fine->f[fine->parity], fine->flags, fine->dims.nx, fine->dims.ny, fine->dims.nz, cf, bad);
    launch_coarse_count(st, cf, bad, nd, coarse->wdist, cnx, cny, cnz, nu0, counts);
    std::vector<uint32_t> hc((size_t)nb);
    hipError_t e = hipMemcpyAsync(hc.data(), counts, sizeof(uint32_t) * (size_t)nb, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      rc = fail(WT_E_HIP, std::string("dataset count: ") + hipGetErrorString(e));
    } else {
      std::vector<uint64_t> ho((size_t)nb);
      uint64_t total = 0;
      for (int64_t b = 0; b < nb; b++) {
        ho[(size_t)b] = (uint64_t)d->count + total;
        total += hc[(size_t)b];
      }
      if (d->count + (int64_t)total > d->cap) {
        rc = fail(WT_E_NOMEM, "dataset capacity exceeded");
      } else {
        e = hipMemcpyAsync(offs, ho.data(), sizeof(uint64_t) * (size_t)nb, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) {
          launch_coarse_samples(st, cf, bad, nd, coarse->wdist, cnx, cny, cnz, nu0, offs, d->data);
          e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = fail(WT_E_HIP, std::string("dataset samples: ") + hipGetErrorString(e));
        else {
          d->count += (int64_t)total;
          if (added) *added = (int64_t)total;
        }
      }
    }
  }
  if (cf) (void)hipFree(cf);
  if (nd) (void)hipFree(nd);
  if (bad) (void)hipFree(bad);
  if (counts) (void)hipFree(counts);
  if (offs) (void)hipFree(offs);
  return rc;
}

namespace {

// Device buffers of one training run.
struct Trainer {
  hipStream_t st = nullptr;
  float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr, *part = nullptr;
  uint32_t* perm = nullptr;
  double* loss = nullptr;
  ~Trainer() {
    if (st) (void)hipStreamSynchronize(st);
    for (void* q : {(void*)p, (void*)m, (void*)v, (void*)g, (void*)part, (void*)perm, (void*)loss})
      if (q) (void)hipFree(q);
    if (st) (void)hipStreamDestroy(st);
  }
  int init(int64_t n, int64_t batch, const float* params) {
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIP_TRY(hipMalloc((void**)&p, sizeof(float) * MLP_NPARAM));
    HIP_TRY(hipMalloc((void**)&m, sizeof(float) * MLP_NPARAM));
    HIP_TRY(hipMalloc((void**)&v, sizeof(float) * MLP_NPARAM));
    HIP_TRY(hipMalloc((void**)&g, sizeof(float) * MLP_NPARAM));
    HIP_TRY(hipMalloc((void**)&part, sizeof(float) * (size_t)grad_partial_floats(batch)));
    HIP_TRY(hipMalloc((void**)&perm, sizeof(uint32_t) * (size_t)std::max<int64_t>(n, 1)));
    HIP_TRY(hipMalloc((void**)&loss, sizeof(double)));
    HIP_TRY(hipMemcpyAsync(p, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(m, 0, sizeof(float) * MLP_NPARAM, st));
    HIP_TRY(hipMemsetAsync(v, 0, sizeof(float) * MLP_NPARAM, st));
    return WT_OK;
  }
};

}  // namespace

int wt_mlp_grad(wt_dataset* d, const wt_mlp* net, int64_t first, int64_t count, float* grad, double* loss) {
  if (!d || !net || !grad || !loss || first < 0 || count < 1 || first + count > d->count) return fail(WT_E_ARG, "bad gradient arguments");
  if (d->count > (int64_t)UINT32_MAX) return fail(WT_E_ARG, "dataset too large");
  Trainer T;
  if (int r = T.init(d->count, count, net->params)) return r;
  std::vector<uint32_t> id((size_t)count);
  std::iota(id.begin(), id.end(), (uint32_t)first);
  HIP_TRY(hipMemcpyAsync(T.perm + first, id.data(), sizeof(uint32_t) * (size_t)count, hipMemcpyHostToDevice, T.st));
  HIP_TRY(hipMemsetAsync(T.loss, 0, sizeof(double), T.st));
  launch_mlp_grad(T.st, d->data, T.perm, first, count, 1.0f / (float)count, T.p, host_norm(net), T.part, T.g, T.loss);
  HIP_TRY(hipGetLastError());
  double l = 0;
  HIP_TRY(hipMemcpyAsync(grad, T.g, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, T.st));
  HIP_TRY(hipMemcpyAsync(&l, T.loss, sizeof(double), hipMemcpyDeviceToHost, T.st));
  HIP_TRY(hipStreamSynchronize(T.st));
  *loss = l / (double)count;
  return WT_OK;
}

int wt_mlp_train(wt_dataset* d, wt_mlp* net, const wt_train_opts* o, double* epoch_loss) {
  if (!d || !net || !o) return fail(WT_E_ARG, "null argument");
  if (o->epochs < 1 || o->batch < 1 || !(o->lr > 0.0)) return fail(WT_E_ARG, "bad training options");
  const int64_t n = o->max_samples > 0 ? std::min<int64_t>(o->max_samples, d->count) : d->count;
  if (n < 1) return fail(WT_E_STATE, "empty dataset");
  if (d->count > (int64_t)UINT32_MAX) return fail(WT_E_ARG, "dataset too large");
  Trainer T;
  if (int r = T.init(n, o->batch, net->params)) return r;
  const MlpNorm norm = host_norm(net);
  std::mt19937_64 rng(o->seed);
  std::vector<uint32_t> all((size_t)d->count);
  std::iota(all.begin(), all.end(), 0u);
  int64_t t = 0;
  for (int ep = 0; ep < o->epochs; ep++) {
    // every epoch draws n samples of the dataset in a fresh order
    std::shuffle(all.begin(), all.end(), rng);
    HIP_TRY(hipMemcpyAsync(T.perm, all.data(), sizeof(uint32_t) * (size_t)n, hipMemcpyHostToDevice, T.st));
    HIP_TRY(hipMemsetAsync(T.loss, 0, sizeof(double), T.st));
    for (int64_t b = 0; b < n; b += o->batch) {
      const int64_t cnt = std::min<int64_t>(o->batch, n - b);
      launch_mlp_grad(T.st, d->data, T.perm, b, cnt, 1.0f / (float)cnt, T.p, norm, T.part, T.g, T.loss);
      launch_adam(T.st, T.p, T.m, T.v, T.g, (float)o->lr, (float)o->beta1, (float)o->beta2, (float)o->eps, ++t);
    }
    HIP_TRY(hipGetLastError());
    double l = 0;
    HIP_TRY(hipMemcpyAsync(&l, T.loss, sizeof(double), hipMemcpyDeviceToHost, T.st));
    HIP_TRY(hipStreamSynchronize(T.st));  // the host order array is reused next epoch
    if (epoch_loss) epoch_loss[ep] = l / (double)n;
  }
  HIP_TRY(hipMemcpyAsync(net->params, T.p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, T.st));
  HIP_TRY(hipStreamSynchronize(T.st));
  return WT_OK;
}

}  // extern "C"
