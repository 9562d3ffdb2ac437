// C-ABI host layer (libwt.so): solver handle, memory ownership, stepping, sampling and readback.
// Mirrors the member functions of SolverGPU (src/solver/SolverGPU.ts); each entry point is
// declared and cited in include/wt.h.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "solver.hpp"
#include "voxel.hpp"

using namespace wt;

namespace {
thread_local std::string g_err;
}  // namespace

int wt::fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

namespace wt {

// SolverGPU.writeParams, SolverGPU.ts:160-187.
void derive(wt_solver* s) {
  KParams& P = s->P;
  const wt_params& p = s->prm;
  P.nx = s->dims.nx; P.ny = s->dims.ny; P.nz = s->dims.nz; P.q = s->q;
  P.n = s->n;
  P.tau0 = (float)(3.0 * p.nu + 0.5);
  P.csmag2 = (float)(p.cs * p.cs);
  P.uin = (float)p.u;
  P.rhoIn = 1.0f;
  P.ground = p.ground;
  P.spongeStart = (float)(p.sponge_start * s->dims.nx);
  P.spongeNu = (float)p.sponge_nu;
  P.uGround = p.ground == WT_GROUND_MOVING ? (float)p.u : 0.0f;
  P.emaAlpha = (float)p.ema_alpha;
  P.uRef = (float)std::max(p.u, 1e-4);
  P.refX = std::min(4, s->dims.nx - 1);
  P.sideMode = p.sides == WT_SIDES_PERIODIC ? 1 : 0;
  P.collision = p.collision == WT_COLL_BGK ? 1 : 0;
  P.tauWall = (float)p.tau_wall;
  P.spongeIn = (float)(p.sponge_in * s->dims.nx);
  P.contactH = 2.0f;
  P.interp = p.interp ? 1 : 0;
  P.closure = p.closure;
  P.invSponge = (float)(1.0 / std::max((double)s->dims.nx - p.sponge_start * s->dims.nx, 1.0));
  P.invSpongeIn = (float)(1.0 / std::max(p.sponge_in * s->dims.nx, 1.0));
  P.smag = (float)(18.0 * std::sqrt(2.0) * p.cs * p.cs);
  P.inu32 = (float)(1.0 / (32.0 * p.nu));
  P.divx.init((uint32_t)s->dims.nx);
  P.divy.init((uint32_t)s->dims.ny);
  P.cpt = (s->cpt == 2 && s->dims.nx % 2 == 0) ? 2 : 1;
}

int check_params(const wt_params* p) {
  if (!(p->u >= 0.0 && p->u < 0.5)) return fail(WT_E_ARG, "params.u must be in [0, 0.5)");
  if (!(p->nu > 0.0)) return fail(WT_E_ARG, "params.nu must be > 0");
  if (!(p->cs >= 0.0)) return fail(WT_E_ARG, "params.cs must be >= 0");
  if (p->ground < 0 || p->ground > 2) return fail(WT_E_ARG, "params.ground must be 0, 1 or 2");
  if (p->sides < 0 || p->sides > 1) return fail(WT_E_ARG, "params.sides must be 0 or 1");
  if (p->precision < 0 || p->precision > 2) return fail(WT_E_ARG, "params.precision must be 0, 1 or 2");
  if (p->collision < 0 || p->collision > 1) return fail(WT_E_ARG, "params.collision must be 0 or 1");
  if (p->closure < 0 || p->closure > 1) return fail(WT_E_ARG, "params.closure must be 0 or 1");
  return WT_OK;
}

int run_init(wt_solver* s) {
  LaunchCtx c = s->ctx();
  launch_init(c, s->f[0]);
  launch_init(c, s->f[1]);
  HIP_TRY(hipGetLastError());
  s->pending_init = false;
  return WT_OK;
}

// Links of the wall cells (depend on flags, distances, wheels and parameters) and the bulk pull
// table (depends on the parameters).
int prep_walls(wt_solver* s) {
  const size_t len = (size_t)pull_table_len(s->q);
  if (!s->pull) HIP_TRY(hipMalloc((void**)&s->pull, len * sizeof(int64_t)));
  // The pull table depends on the grid and the boundary modes only: the per-frame parameter ramps
  // (speed, viscosity, averaging rate) leave it unchanged, and then nothing is uploaded and the
  // call does not wait for the stream. A changed table is uploaded in stream order (after the steps
  // already queued, which read the old one) and waited for, since the host copy is its source.
  std::vector<int64_t> t(len, 0);
  pull_table(s->P, t.data());
  if (t != s->h_pull) {
    hipError_t e = hipMemcpyAsync(s->pull, t.data(), len * sizeof(int64_t), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
      s->h_pull.clear();  // the device table is not known to match any host copy
      return fail(WT_E_HIP, std::string("pull table upload: ") + hipGetErrorString(e));
    }
    s->h_pull.swap(t);
  }
  if (!s->wlink) return WT_OK;
  LaunchCtx c = s->ctx();
  launch_wallprep(c, s->wlist, s->wmask, s->wlink, s->wmeta, s->nwall, s->wcap);
  HIP_TRY(hipGetLastError());
  return WT_OK;
}

int reclassify(wt_solver* s) {
  s->wdist_valid = false;  // flags changed
  HIP_TRY(hipMemsetAsync(s->counts, 0, 4 * sizeof(unsigned long long), s->stream));
  LaunchCtx c = s->ctx();
  launch_classify(c, s->wlist, s->wmask, s->counts);
  HIP_TRY(hipGetLastError());
  unsigned long long h[4];
  HIP_TRY(hipMemcpyAsync(h, s->counts, sizeof(h), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  s->nwall = (int64_t)h[0];
  s->nfast = s->n - (int64_t)h[0] - (int64_t)h[2];  // every cell is listed, bulk solid or bulk fluid
  s->nsolid = (int64_t)h[2];
  // the wall list in cell order (the collection appends wave by wave in any order), so that each
  // group of wall blocks covers a short range of rows and can be placed beside them (build_roles)
  s->h_wrows.clear();
  if (s->nwall > 0) {
    const size_t nw = (size_t)s->nwall;
    std::vector<uint32_t> cells(nw), masks(nw), order(nw);
    HIP_TRY(hipMemcpyAsync(cells.data(), s->wlist, 4 * nw, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(masks.data(), s->wmask, 4 * nw, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (size_t i = 0; i < nw; i++) order[i] = (uint32_t)i;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cells[a] < cells[b]; });
    std::vector<uint32_t> sc(nw), sm(nw);
    for (size_t i = 0; i < nw; i++) {
      sc[i] = cells[order[i]];
      sm[i] = masks[order[i]];
    }
    HIP_TRY(hipMemcpyAsync(s->wlist, sc.data(), 4 * nw, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->wmask, sm.data(), 4 * nw, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (size_t i = 0; i < nw; i += WALL_GROUP_CELLS) s->h_wrows.push_back(sc[i] / (uint32_t)s->dims.nx);
  }
  s->wall_version++;
  if (s->nwall > s->wcap || !s->wlink) {
    if (s->wlink) HIP_TRY(hipFree(s->wlink));
    if (s->wmeta) HIP_TRY(hipFree(s->wmeta));
    s->wlink = nullptr;
    s->wmeta = nullptr;
    s->wcap = 0;
    const int64_t cap = s->nwall + s->nwall / 4 + 1024;
    if (hipMalloc((void**)&s->wlink, sizeof(WallLink) * (size_t)(s->q - 1) * (size_t)cap) != hipSuccess ||
        hipMalloc((void**)&s->wmeta, sizeof(uint32_t) * (size_t)cap) != hipSuccess)
      return fail(WT_E_NOMEM, "wall link allocation failed");
    s->wcap = cap;
  }
  return prep_walls(s);
}

int ensure_roles(wt_solver* s, StepArgs& A) {
  const uint64_t key[6] = {A.nwb, A.nbb, A.xk, A.wpr, A.nwaves, s->wall_version};
  if (!s->role || std::memcmp(key, s->role_key, sizeof(key)) != 0) {
    const std::vector<uint32_t> tab = build_roles(A, s->h_wrows, s->dims.ny, s->dims.nz);
    if (tab.size() > s->role_cap) {
      if (s->role) HIP_TRY(hipFree(s->role));
      s->role = nullptr;
      s->role_cap = 0;
      HIP_TRY(hipMalloc((void**)&s->role, 4 * tab.size()));
      s->role_cap = tab.size();
    }
    HIP_TRY(hipMemcpyAsync(s->role, tab.data(), 4 * tab.size(), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    std::memcpy(s->role_key, key, sizeof(key));
  }
  A.role = s->role;
  return WT_OK;
}

int do_steps(wt_solver* s, int steps) {
  if (s->pending_init) {
    int r = run_init(s);
    if (r) return r;
  }
  if (s->P.closure == 1 && s->mlp && !s->wdist_valid) {
    if (int r = ensure_wdist(s)) return r;
  }
  LaunchCtx c = s->ctx();
  StepArgs A = step_args(c, s->wlist, s->wmeta, s->wlink, s->nwall, s->wcap, s->item_chunk);
  if (int r = ensure_roles(s, A)) return r;
  // next free event pair, covering `launches` step launches
  auto pair = [&](int64_t launches, hipEvent_t*& e) -> int {
    if (s->ev_used + 2 > s->ev.size()) {
      const size_t old = s->ev.size();
      s->ev.resize(std::max<size_t>(64, old * 2));
      for (size_t i = old; i < s->ev.size(); i++) HIP_TRY(hipEventCreate(&s->ev[i]));
    }
    e = &s->ev[s->ev_used];
    s->ev_used += 2;
    s->ev_launches.push_back(launches);
    return WT_OK;
  };
  hipEvent_t* span = nullptr;
  if (s->prof == WT_PROFILE_SPAN && steps > 0) {
    if (int r = pair(steps, span)) return r;
    HIP_TRY(hipEventRecord(span[0], s->stream));
  }
  for (int k = 0; k < steps; k++) {
    const void* fin = s->f[s->parity];
    void* fout = s->f[s->parity ^ 1];
    if (s->prof == WT_PROFILE_LAUNCH) {
      hipEvent_t* e = nullptr;
      if (int r = pair(1, e)) return r;
      launch_step(c, fin, fout, A, e[0], e[1]);
    } else {
      launch_step(c, fin, fout, A);
    }
    s->parity ^= 1;
  }
  if (span) HIP_TRY(hipEventRecord(span[1], s->stream));
  HIP_TRY(hipGetLastError());
  s->step += steps;
  return WT_OK;
}

int do_sample(wt_solver* s) {
  if (s->pending_init) {
    int r = run_init(s);
    if (r) return r;
  }
  LaunchCtx c = s->ctx();
  launch_macro(c, s->f[s->parity], s->vel, s->meanA, s->meanB, s->partials, s->result);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(s->h_result, s->result, 8 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return WT_OK;
}

// SolverGPU.ts:302-308 (and the same decode in flushReads, :270-279)
void decode_result(const double* d, int64_t step, wt_stats* out) {
  out->fx = d[0]; out->fy = d[1]; out->fz = d[2];
  out->max_u = std::sqrt(std::max(d[3], 0.0));
  out->rho_ref = d[5] > 0 ? d[4] / d[5] : 1.0;
  out->unstable = (d[6] > 0 || !std::isfinite(d[0]) || d[3] > 0.25) ? 1 : 0;
  out->pad = 0;
  out->step = step;
}

void decode_stats(const wt_solver* s, wt_stats* out) { decode_result(s->h_result, s->step, out); }

int64_t storage_bytes(const wt_solver* s) { return (int64_t)s->q * s->n * (s->prec == WT_PREC_F32 ? 4 : 2); }

}  // namespace wt

// A batch of independent domains stepped by one launch per step (k_step_batch).
struct wt_batch {
  std::vector<wt_solver*> m;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  DomainDesc* d_desc = nullptr;
  uint32_t* d_map = nullptr;
  int64_t map_cap = 0;
  std::vector<DomainDesc> h_desc;
  uint32_t blocks = 0;
  int q = 0, prec = 0, cpt = 0;
  bool reg = true;
  int chunk = 100;      // steps per captured graph (even)
  bool use_graph = true;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};  // per starting parity
  hipGraph_t graph[2] = {nullptr, nullptr};
};

namespace {

void batch_drop_graph(wt_batch* b) {
  for (int p = 0; p < 2; p++) {
    if (b->gexec[p]) (void)hipGraphExecDestroy(b->gexec[p]);
    if (b->graph[p]) (void)hipGraphDestroy(b->graph[p]);
    b->gexec[p] = nullptr;
    b->graph[p] = nullptr;
  }
}

// The batched kernel variant is fixed by the batch's lattice, storage, cells per thread and
// collision; a member whose state no longer matches it (retuned, switched collision, learned
// closure attached: k_step_batch has no network path) cannot be stepped by it.
int batch_member_check(const wt_solver* s, int q, int prec, int cpt, bool reg) {
  if (s->q != q || s->prec != prec) return fail(WT_E_ARG, "batch members must share lattice and storage precision");
  if (cells_per_thread(s->P) != cpt)
    return fail(WT_E_STATE, "a batch member's cells per thread differs from the batch's (retuned after the batch was created)");
  if ((s->P.collision != 1) != reg)
    return fail(WT_E_STATE, "a batch member's collision operator differs from the batch's (changed after the batch was created)");
  if (s->P.closure == WT_CLOSURE_MLP && s->mlp)
    return fail(WT_E_STATE, "a batch member uses the learned closure, which the batched step does not run");
  return WT_OK;
}

// Descriptor table and block -> domain map for the members' current state; uploads and drops the
// captured graph when anything changed.
int batch_prepare(wt_batch* b) {
  std::vector<DomainDesc> desc(b->m.size());
  uint32_t blocks = 0;
  for (size_t k = 0; k < b->m.size(); k++) {
    if (int r = batch_member_check(b->m[k], b->q, b->prec, b->cpt, b->reg)) return r;
  }
  for (size_t k = 0; k < b->m.size(); k++) {
    wt_solver* s = b->m[k];
    LaunchCtx c = s->ctx();
    DomainDesc& d = desc[k];
    std::memset(&d, 0, sizeof(d));
    d.P = s->P;
    d.A = step_args(c, s->wlist, s->wmeta, s->wlink, s->nwall, s->wcap, s->item_chunk);
    if (int r = ensure_roles(s, d.A)) return r;
    d.f[0] = s->f[0];
    d.f[1] = s->f[1];
    d.kind = s->kind;
    d.first = blocks;
    blocks += d.A.nwb + d.A.nbb;  // multiples of 8: every domain's blocks keep their XCD
  }
  const bool same = desc.size() == b->h_desc.size() &&
                    std::memcmp(desc.data(), b->h_desc.data(), desc.size() * sizeof(DomainDesc)) == 0;
  if (same && b->d_desc) return WT_OK;
  batch_drop_graph(b);
  std::vector<uint32_t> map(blocks);
  for (size_t k = 0; k < b->m.size(); k++)
    for (uint32_t i = 0; i < desc[k].A.nwb + desc[k].A.nbb; i++) map[desc[k].first + i] = (uint32_t)k;
  if (!b->d_desc) HIP_TRY(hipMalloc((void**)&b->d_desc, sizeof(DomainDesc) * b->m.size()));
  if ((int64_t)map.size() > b->map_cap) {
    if (b->d_map) HIP_TRY(hipFree(b->d_map));
    b->d_map = nullptr;
    HIP_TRY(hipMalloc((void**)&b->d_map, sizeof(uint32_t) * map.size()));
    b->map_cap = (int64_t)map.size();
  }
  HIP_TRY(hipMemcpyAsync(b->d_desc, desc.data(), sizeof(DomainDesc) * desc.size(), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipMemcpyAsync(b->d_map, map.data(), sizeof(uint32_t) * map.size(), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->h_desc = desc;
  b->blocks = blocks;
  return WT_OK;
}

void batch_launch(wt_batch* b, int parity) {
  launch_step_batch(b->stream, b->q, b->prec, b->cpt, b->reg, b->d_desc, b->d_map, b->blocks, parity);
}

}  // namespace

extern "C" {

const char* wt_last_error(void) { return g_err.c_str(); }

int64_t wt_bytes_for(const wt_dims* d) {
  if (!d) return 0;
  const int64_t n = (int64_t)d->nx * d->ny * d->nz;
  const int q = d->nz > 1 ? 19 : 9;
  return n * (q * 8 + 4 + 20 + 32);
}

int wt_create(const wt_dims* dims, const wt_params* params, wt_solver** out) {
  if (!dims || !params || !out) return fail(WT_E_ARG, "null argument");
  *out = nullptr;
  if (dims->nx < 2 || dims->ny < 2 || dims->nz < 1) return fail(WT_E_ARG, "dims must be nx >= 2, ny >= 2, nz >= 1");
  const int64_t n = (int64_t)dims->nx * dims->ny * dims->nz;
  if (n >= (1ll << 31)) return fail(WT_E_ARG, "grid too large: cell count must be below 2^31");
  if (int r = check_params(params)) return r;
  int ndev = 0, dev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(WT_E_NODEV, "no HIP device available");
  HIP_TRY(hipGetDevice(&dev));
  auto* s = new wt_solver();
  s->device = dev;
  s->dims = *dims;
  s->n = n;
  s->q = dims->nz > 1 ? 19 : 9;
  s->prm = *params;
  // half-width storage needs an even nx (pairs of cells per 32-bit word), else FP32, as SolverGPU.ts:69
  s->prec = (params->precision != WT_PREC_F32 && dims->nx % 2 == 0) ? params->precision : WT_PREC_F32;
  s->prm.precision = s->prec;
  derive(s);
  const int64_t sbytes = s->prec == WT_PREC_F32 ? 4 : 2;
  const int64_t fbytes = (int64_t)s->q * n * sbytes;
  const int64_t nrecords = macro_records(n);
  auto alloc = [&](void** p, int64_t bytes) -> bool {
    if (hipMalloc(p, (size_t)std::max<int64_t>(bytes, 256)) != hipSuccess) return false;
    s->bytes += bytes;
    return true;
  };
  // f[] gets one extra word of padding so pair loads at the very last cell stay in bounds
  bool ok = alloc(&s->f[0], fbytes + 64) && alloc(&s->f[1], fbytes + 64) && alloc((void**)&s->flags, 4 * n) &&
            alloc((void**)&s->sdf, 4 * n) && alloc((void**)&s->kind, n + 64) && alloc((void**)&s->wlist, 4 * (n + 1)) && alloc((void**)&s->wmask, 4 * (n + 1)) &&
            alloc((void**)&s->counts, 64) && alloc((void**)&s->vel, 8 * n) && alloc((void**)&s->meanA, 16 * n) &&
            alloc((void**)&s->meanB, 4 * n) && alloc((void**)&s->partials, 64 * nrecords) && alloc((void**)&s->result, 64);
  if (!ok) {
    wt_destroy(s);
    return fail(WT_E_NOMEM, "device allocation failed");
  }
  if (hipHostMalloc((void**)&s->h_result, 64, hipHostMallocDefault) != hipSuccess) {
    wt_destroy(s);
    return fail(WT_E_NOMEM, "pinned host allocation failed");
  }
  std::memset(s->h_result, 0, 64);  // wt_sample_raw before the first sample reads zeros
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    wt_destroy(s);
    return fail(WT_E_HIP, "hipStreamCreate failed");
  }
  {
    hipError_t e = hipMemsetAsync(s->flags, 0, 4 * n, s->stream);
    if (e == hipSuccess) e = hipMemsetAsync(s->sdf, 0, 4 * n, s->stream);
    if (e == hipSuccess) e = hipMemsetAsync(s->f[0], 0, fbytes + 64, s->stream);
    if (e == hipSuccess) e = hipMemsetAsync(s->f[1], 0, fbytes + 64, s->stream);
    if (e == hipSuccess) e = hipMemsetAsync(s->kind, 0, n + 64, s->stream);
    if (e != hipSuccess) {
      wt_destroy(s);
      return fail(WT_E_HIP, std::string("wt_create: ") + hipGetErrorString(e));
    }
  }
  std::memset(&s->W, 0, sizeof(s->W));
  if (int r = reclassify(s)) { wt_destroy(s); return r; }
  if (int r = wt_reset(s)) { wt_destroy(s); return r; }
  *out = s;
  return WT_OK;
}

int wt_destroy(wt_solver* s) {
  if (!s) return WT_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (auto e : s->ev) (void)hipEventDestroy(e);
  for (auto e : s->fence_ev)
    if (e) (void)hipEventDestroy(e);
  void* ptrs[] = {s->f[0], s->f[1], s->flags, s->sdf, s->kind, s->wlist, s->wmask, s->wlink, s->wmeta, s->wdist, s->mlp, s->counts, s->vel, s->meanA, s->meanB, s->partials, s->result, s->vort, s->stat, s->meanh, s->pull, s->role};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (s->h_result) (void)hipHostFree(s->h_result);
  for (auto& r : s->reads) {
    if (r.ev) (void)hipEventDestroy(r.ev);
    if (r.host) (void)hipHostFree(r.host);
  }
  for (auto* r : s->samples) {
    if (r->ev) (void)hipEventDestroy(r->ev);
    if (r->host) (void)hipHostFree(r->host);
    delete r;
  }
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return WT_OK;
}

int wt_info_get(const wt_solver* s, wt_info* out) {
  if (!s || !out) return fail(WT_E_ARG, "null argument");
  out->dims = s->dims;
  out->n = s->n;
  out->q = s->q;
  out->precision = s->prec;
  out->step = s->step;
  out->wall_cells = s->nwall;
  out->device_bytes = s->bytes;
  return WT_OK;
}

int wt_set_params(wt_solver* s, const wt_params* p, uint32_t mask) {
  if (!s || !p) return fail(WT_E_ARG, "null argument");
  wt_params np = s->prm;
  if (mask & WT_P_U) np.u = p->u;
  if (mask & WT_P_NU) np.nu = p->nu;
  if (mask & WT_P_CS) np.cs = p->cs;
  if (mask & WT_P_GROUND) np.ground = p->ground;
  if (mask & WT_P_SIDES) np.sides = p->sides;
  if (mask & WT_P_SPONGE_NU) np.sponge_nu = p->sponge_nu;
  if (mask & WT_P_SPONGE_START) np.sponge_start = p->sponge_start;
  if (mask & WT_P_SPONGE_IN) np.sponge_in = p->sponge_in;
  if (mask & WT_P_TAU_WALL) np.tau_wall = p->tau_wall;
  if (mask & WT_P_INTERP) np.interp = p->interp;
  if (mask & WT_P_COLLISION) np.collision = p->collision;
  if (mask & WT_P_CLOSURE) np.closure = p->closure;
  if (mask & WT_P_EMA_ALPHA) np.ema_alpha = p->ema_alpha;
  if (int r = check_params(&np)) return r;
  np.precision = s->prec;  // fixed at construction
  s->prm = np;
  derive(s);
  return prep_walls(s);  // the wall links hold parameter-dependent terms
}

int wt_get_params(const wt_solver* s, wt_params* out) {
  if (!s || !out) return fail(WT_E_ARG, "null argument");
  *out = s->prm;
  return WT_OK;
}

double wt_tau0(const wt_solver* s) { return s ? 3.0 * s->prm.nu + 0.5 : NAN; }

int wt_reset(wt_solver* s) {
  if (!s) return fail(WT_E_ARG, "null argument");
  s->pending_init = true;
  s->step = 0;
  return wt_clear_stats(s);
}

int wt_clear_stats(wt_solver* s) {
  if (!s) return fail(WT_E_ARG, "null argument");
  LaunchCtx c = s->ctx();
  launch_clear_stats(c, s->meanA, s->meanB);
  HIP_TRY(hipGetLastError());
  return WT_OK;
}

int wt_upload_flags(wt_solver* s, const uint32_t* flags) {
  if (!s || !flags) return fail(WT_E_ARG, "null argument");
  HIP_TRY(hipMemcpyAsync(s->flags, flags, 4 * s->n, hipMemcpyHostToDevice, s->stream));
  return reclassify(s);
}

int wt_read_flags(wt_solver* s, uint32_t* out) {
  if (!s || !out) return fail(WT_E_ARG, "null argument");
  HIP_TRY(hipMemcpyAsync(out, s->flags, 4 * s->n, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return WT_OK;
}

int wt_upload_sdf(wt_solver* s, const float* d) {
  if (!s || !d) return fail(WT_E_ARG, "null argument");
  std::vector<uint32_t> enc((size_t)s->n);
  for (int64_t i = 0; i < s->n; i++) {
    if (!(d[i] >= 0.0f)) { enc[i] = 0u; continue; }
    uint32_t b;
    std::memcpy(&b, &d[i], 4);
    enc[i] = ~b;
  }
  HIP_TRY(hipMemcpyAsync(s->sdf, enc.data(), 4 * s->n, hipMemcpyHostToDevice, s->stream));
  if (int r = prep_walls(s)) return r;
  HIP_TRY(hipStreamSynchronize(s->stream));
  return WT_OK;
}

int wt_read_sdf(wt_solver* s, float* d) {
  if (!s || !d) return fail(WT_E_ARG, "null argument");
  std::vector<uint32_t> enc((size_t)s->n);
  HIP_TRY(hipMemcpyAsync(enc.data(), s->sdf, 4 * s->n, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (int64_t i = 0; i < s->n; i++) {
    if (enc[i] == 0u) { d[i] = -1.0f; continue; }
    const uint32_t b = ~enc[i];
    std::memcpy(&d[i], &b, 4);
  }
  return WT_OK;
}

int wt_set_wheels(wt_solver* s, const float* w, int32_t count) {
  if (!s) return fail(WT_E_ARG, "null argument");
  if (count < 0 || (count > 0 && !w)) return fail(WT_E_ARG, "bad wheel list");
  // the wheel table lives in the kernel argument block; wait for in-flight steps that use the old one
  HIP_TRY(hipStreamSynchronize(s->stream));
  std::memset(&s->W, 0, sizeof(s->W));
  const int m = std::min(count, 16);
  s->W.count = m;
  for (int k = 0; k < m; k++) {
    const float* e = w + 8 * k;
    s->W.w[2 * k] = make_float4(e[0], e[1], e[2], 1.0f / std::max(e[6], 1e-3f));
    s->W.w[2 * k + 1] = make_float4(e[3], e[4], e[5], e[7]);
  }
  return prep_walls(s);  // the wall links hold the wheel velocities
}

int wt_step(wt_solver* s, int32_t steps) {
  if (!s) return fail(WT_E_ARG, "null argument");
  if (steps < 0) return fail(WT_E_ARG, "steps must be >= 0");
  return do_steps(s, steps);
}

int wt_step_and_sample(wt_solver* s, int32_t steps, wt_stats* out) {
  if (!s || !out) return fail(WT_E_ARG, "null argument");
  if (steps < 0) return fail(WT_E_ARG, "steps must be >= 0");
  if (int r = do_steps(s, steps)) return r;
  if (int r = do_sample(s)) return r;
  decode_stats(s, out);
  return WT_OK;
}

int wt_step_and_sample_async(wt_solver* s, int32_t steps, int64_t* ticket) {
  if (!s || !ticket) return fail(WT_E_ARG, "null argument");
  if (steps < 0) return fail(WT_E_ARG, "steps must be >= 0");
  if (int r = do_steps(s, steps)) return r;  // runs a pending init first
  LaunchCtx c = s->ctx();
  launch_macro(c, s->f[s->parity], s->vel, s->meanA, s->meanB, s->partials, s->result);
  HIP_TRY(hipGetLastError());
  std::lock_guard<std::mutex> g(s->read_mu);
  size_t idx = 0;
  while (idx < s->samples.size() && s->samples[idx]->busy) idx++;
  if (idx == s->samples.size()) {
    auto* r = new wt_solver::ReadSlot();
    if (hipHostMalloc((void**)&r->host, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&r->ev, hipEventDisableTiming) != hipSuccess) {
      if (r->host) (void)hipHostFree(r->host);
      delete r;
      return fail(WT_E_NOMEM, "sample slot allocation failed");
    }
    s->samples.push_back(r);
  }
  wt_solver::ReadSlot* r = s->samples[idx];
  HIP_TRY(hipMemcpyAsync(r->host, s->result, 8 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipEventRecord(r->ev, s->stream));
  r->busy = true;
  r->step = s->step;
  r->gen++;
  *ticket = ((int64_t)idx << 32) | (int64_t)r->gen;
  return WT_OK;
}

int wt_sample_wait(wt_solver* s, int64_t ticket, wt_stats* out) {
  if (!s || !out) return fail(WT_E_ARG, "null argument");
  const size_t idx = (size_t)((uint64_t)ticket >> 32);
  const uint32_t gen = (uint32_t)((uint64_t)ticket & 0xffffffffu);
  hipEvent_t ev = nullptr;
  wt_solver::ReadSlot* r = nullptr;
  {
    std::lock_guard<std::mutex> g(s->read_mu);
    if (idx >= s->samples.size() || !s->samples[idx]->busy || s->samples[idx]->gen != gen)
      return fail(WT_E_ARG, "unknown or already collected sample ticket");
    r = s->samples[idx];
    ev = r->ev;
  }
  // the wait touches only the ticket's slot: other calls on the handle may run meanwhile
  const hipError_t e = hipEventSynchronize(ev);
  std::lock_guard<std::mutex> g(s->read_mu);
  if (e == hipSuccess) decode_result(r->host, r->step, out);
  r->busy = false;
  if (e != hipSuccess) return fail(WT_E_HIP, std::string("wt_sample_wait: ") + hipGetErrorString(e));
  return WT_OK;
}

int wt_sample_raw(wt_solver* s, double* out8) {
  if (!s || !out8) return fail(WT_E_ARG, "null argument");
  std::memcpy(out8, s->h_result, 8 * sizeof(double));
  return WT_OK;
}

int wt_sync(wt_solver* s) {
  if (!s) return fail(WT_E_ARG, "null argument");
  HIP_TRY(hipStreamSynchronize(s->stream));
  return WT_OK;
}

int wt_fence(wt_solver* s, int64_t* ticket) {
  if (!s || !ticket) return fail(WT_E_ARG, "null argument");
  const int64_t t = s->fence_seq + 1;
  const int k = (int)(t % wt_solver::FENCES);
  if (!s->fence_ev[k]) HIP_TRY(hipEventCreateWithFlags(&s->fence_ev[k], hipEventDisableTiming));
  HIP_TRY(hipEventRecord(s->fence_ev[k], s->stream));
  s->fence_seq = t;
  s->fence_ticket[k] = t;
  *ticket = t;
  return WT_OK;
}

int wt_fence_done(wt_solver* s, int64_t ticket, int32_t* done) {
  if (!s || !done) return fail(WT_E_ARG, "null argument");
  if (ticket < 1 || ticket > s->fence_seq) return fail(WT_E_ARG, "wt_fence_done: not a ticket of this handle");
  *done = 0;
  if (ticket <= s->fence_done) {
    *done = 1;
    return WT_OK;
  }
  // the ticket's slot holds it or a newer fence (reused after FENCES newer ones): passing that one
  // passes this one too; not having passed it yet says nothing, so 0 until it has
  const int k = (int)(ticket % wt_solver::FENCES);
  const hipError_t q = hipEventQuery(s->fence_ev[k]);
  if (q == hipErrorNotReady) return WT_OK;
  if (q != hipSuccess) return fail(WT_E_HIP, std::string("wt_fence_done: ") + hipGetErrorString(q));
  s->fence_done = std::max(s->fence_done, s->fence_ticket[k]);
  *done = 1;
  return WT_OK;
}

int wt_fence_wait(wt_solver* s, int64_t ticket) {
  if (!s) return fail(WT_E_ARG, "null argument");
  if (ticket < 1 || ticket > s->fence_seq) return fail(WT_E_ARG, "wt_fence_wait: not a ticket of this handle");
  if (ticket <= s->fence_done) return WT_OK;
  const int k = (int)(ticket % wt_solver::FENCES);
  HIP_TRY(hipEventSynchronize(s->fence_ev[k]));
  s->fence_done = std::max(s->fence_done, s->fence_ticket[k]);
  return WT_OK;
}

int wt_read_populations(wt_solver* s, float* out) {
  if (!s || !out) return fail(WT_E_ARG, "null argument");
  if (s->pending_init) {
    if (int r = run_init(s)) return r;
  }
  const int64_t cnt = (int64_t)s->q * s->n;
  float* tmp = nullptr;
  HIP_TRY(hipMallocAsync((void**)&tmp, 4 * cnt, s->stream));
  LaunchCtx c = s->ctx();
  launch_unpack(c, s->f[s->parity], tmp);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, tmp, 4 * cnt, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipFreeAsync(tmp, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return WT_OK;
}

int wt_read_population_planes(wt_solver* s, int32_t z0, int32_t count, float* out) {
  if (!s || !out) return fail(WT_E_ARG, "null argument");
  if (count < 1 || z0 < 0 || (int64_t)z0 + count > s->dims.nz) return fail(WT_E_ARG, "planes must lie in [0, nz)");
  if (s->pending_init) {
    if (int r = run_init(s)) return r;
  }
  const int64_t cnt = (int64_t)s->q * s->dims.nx * s->dims.ny * count;
  float* tmp = nullptr;
  HIP_TRY(hipMallocAsync((void**)&tmp, 4 * cnt, s->stream));
  LaunchCtx c = s->ctx();
  launch_unpack_planes(c, s->f[s->parity], z0, count, tmp);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, 4 * cnt, hipMemcpyDeviceToHost, s->stream);
  const hipError_t e2 = hipFreeAsync(tmp, s->stream);
  if (e == hipSuccess) e = e2;
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) return fail(WT_E_HIP, std::string("wt_read_population_planes: ") + hipGetErrorString(e));
  return WT_OK;
}

int wt_upload_populations(wt_solver* s, const float* f) {
  if (!s || !f) return fail(WT_E_ARG, "null argument");
  if (s->pending_init) {
    if (int r = run_init(s)) return r;
  }
  const int64_t cnt = (int64_t)s->q * s->n;
  float* tmp = nullptr;
  HIP_TRY(hipMallocAsync((void**)&tmp, 4 * cnt, s->stream));
  HIP_TRY(hipMemcpyAsync(tmp, f, 4 * cnt, hipMemcpyHostToDevice, s->stream));
  LaunchCtx c = s->ctx();
  launch_pack(c, tmp, s->f[s->parity]);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipFreeAsync(tmp, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return WT_OK;
}

int wt_read_macro(wt_solver* s, float* vel4, float* mean_a4, float* mean_b) {
  if (!s) return fail(WT_E_ARG, "null argument");
  if (vel4) {
    std::vector<uint16_t> h((size_t)(4 * s->n));
    HIP_TRY(hipMemcpyAsync(h.data(), s->vel, 8 * s->n, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (int64_t i = 0; i < 4 * s->n; i++) {
      // IEEE half -> float on the host
      const uint32_t hv = h[i];
      const uint32_t sign = (hv & 0x8000u) << 16, ex = (hv >> 10) & 0x1f, man = hv & 0x3ff;
      float v;
      if (ex == 0) v = std::ldexp((float)man, -24);
      else if (ex == 31) v = man ? NAN : INFINITY;
      else v = std::ldexp((float)(man | 0x400), (int)ex - 25);
      uint32_t b;
      std::memcpy(&b, &v, 4);
      b |= sign;
      std::memcpy(&vel4[i], &b, 4);
    }
  }
  if (mean_a4) HIP_TRY(hipMemcpyAsync(mean_a4, s->meanA, 16 * s->n, hipMemcpyDeviceToHost, s->stream));
  if (mean_b) HIP_TRY(hipMemcpyAsync(mean_b, s->meanB, 4 * s->n, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return WT_OK;
}

// ---- derived fields and export -------------------------------------------------------------

namespace wt {
// The three derived-field buffers start zeroed (the reference's textures before the first pass).
static int ensure_derived(wt_solver* s) {
  if (s->vort) return WT_OK;
  const size_t b = 8 * (size_t)s->n;
  HIP_TRY(hipMalloc(&s->vort, b));
  HIP_TRY(hipMalloc(&s->stat, b));
  HIP_TRY(hipMalloc(&s->meanh, b));
  HIP_TRY(hipMemsetAsync(s->vort, 0, b, s->stream));
  HIP_TRY(hipMemsetAsync(s->stat, 0, b, s->stream));
  HIP_TRY(hipMemsetAsync(s->meanh, 0, b, s->stream));
  return WT_OK;
}
}  // namespace wt

int wt_derive(wt_solver* s) {
  if (!s) return fail(WT_E_ARG, "null argument");
  if (int r = wt::ensure_derived(s)) return r;
  wt::launch_derived(s->ctx(), s->vel, s->meanA, s->meanB, s->vort, s->stat, s->meanh);
  HIP_TRY(hipGetLastError());
  return WT_OK;
}

int wt_encode_post(wt_solver* s, int32_t derived, int32_t* queued) {
  if (!s) return fail(WT_E_ARG, "null argument");
  if (queued) *queued = 0;
  if (s->pending_init) {
    if (int r = run_init(s)) return r;
  }
  if (derived) {
    if (int r = wt::ensure_derived(s)) return r;
  }
  LaunchCtx c = s->ctx();
  launch_macro(c, s->f[s->parity], s->vel, s->meanA, s->meanB, s->partials, s->result);
  if (derived) launch_derived(c, s->vel, s->meanA, s->meanB, s->vort, s->stat, s->meanh);
  HIP_TRY(hipGetLastError());
  if (!derived) return WT_OK;  // the reference reads back only with the derived pass (SolverGPU.ts:247)
  if (s->reads.empty()) {
    s->reads.resize(4);
    for (auto& r : s->reads) {
      if (hipHostMalloc((void**)&r.host, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return fail(WT_E_NOMEM, "pinned read slot allocation failed");
      std::memset(r.host, 0, 8 * sizeof(double));
      HIP_TRY(hipEventCreateWithFlags(&r.ev, hipEventDisableTiming));
    }
  }
  for (auto& r : s->reads) {
    if (r.busy) continue;
    HIP_TRY(hipMemcpyAsync(r.host, s->result, 8 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipEventRecord(r.ev, s->stream));
    r.busy = true;
    r.step = s->step;
    r.seq = ++s->read_seq;
    if (queued) *queued = 1;
    return WT_OK;
  }
  return WT_OK;  // every slot busy: this post is not read back (SolverGPU.ts:249-250)
}

int wt_flush_reads(wt_solver* s, wt_stats* out, int32_t cap, int32_t* count, int32_t* pending) {
  if (!s || !count || (cap > 0 && !out) || cap < 0) return fail(WT_E_ARG, "bad flush arguments");
  *count = 0;
  // completed reads in the order they were posted (one stream: they complete in that order)
  std::vector<wt_solver::ReadSlot*> order;
  for (auto& r : s->reads)
    if (r.busy) order.push_back(&r);
  std::sort(order.begin(), order.end(), [](const wt_solver::ReadSlot* a, const wt_solver::ReadSlot* b) { return a->seq < b->seq; });
  int32_t left = 0;
  for (auto* r : order) {
    if (*count >= cap) { left++; continue; }
    const hipError_t q = hipEventQuery(r->ev);
    if (q == hipErrorNotReady) { left++; continue; }
    if (q != hipSuccess) return fail(WT_E_HIP, std::string("wt_flush_reads: ") + hipGetErrorString(q));
    decode_result(r->host, r->step, &out[*count]);
    (*count)++;
    r->busy = false;
  }
  if (pending) *pending = left;
  return WT_OK;
}

int wt_read_derived(wt_solver* s, float* vort4, float* stat4, float* mean4) {
  if (!s) return fail(WT_E_ARG, "null argument");
  if (int r = wt::ensure_derived(s)) return r;
  float* tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, 16 * (size_t)s->n));
  const ushort4* src[3] = {s->vort, s->stat, s->meanh};
  float* dst[3] = {vort4, stat4, mean4};
  int rc = WT_OK;
  for (int k = 0; k < 3 && rc == WT_OK; k++) {
    if (!dst[k]) continue;
    wt::launch_unpack_h4(s->stream, src[k], s->n, tmp);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(dst[k], tmp, 16 * (size_t)s->n, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) rc = fail(WT_E_HIP, std::string("wt_read_derived: ") + hipGetErrorString(e));
  }
  (void)hipFree(tmp);
  return rc;
}

int wt_export_dims(const wt_solver* s, int32_t stride, int32_t* out_dims) {
  if (!s || !out_dims) return fail(WT_E_ARG, "null argument");
  if (stride < 1) return fail(WT_E_ARG, "stride must be >= 1");
  out_dims[0] = (int32_t)((s->dims.nx + stride - 1) / stride);
  out_dims[1] = (int32_t)((s->dims.ny + stride - 1) / stride);
  out_dims[2] = (int32_t)((s->dims.nz + stride - 1) / stride);
  return WT_OK;
}

int wt_export_fields(wt_solver* s, int32_t stride, float* vel4, float* mean4, float* vort4, int32_t* out_dims) {
  if (!s) return fail(WT_E_ARG, "null argument");
  int32_t o[3];
  if (int r = wt_export_dims(s, stride, o)) return r;
  if (out_dims) for (int k = 0; k < 3; k++) out_dims[k] = o[k];
  if (!vel4 && !mean4 && !vort4) return WT_OK;
  if (int r = wt::ensure_derived(s)) return r;
  const int64_t nout = (int64_t)o[0] * o[1] * o[2];
  float* tmp = nullptr;
  HIP_TRY(hipMalloc(&tmp, 48 * (size_t)nout));
  const int oi[3] = {o[0], o[1], o[2]};
  wt::launch_export(s->stream, s->vel, s->meanh, s->vort, s->dims.nx, s->dims.ny, s->dims.nz, stride, oi, tmp, tmp + 4 * nout,
                    tmp + 8 * nout);
  hipError_t e = hipGetLastError();
  float* dst[3] = {vel4, mean4, vort4};
  for (int k = 0; k < 3 && e == hipSuccess; k++)
    if (dst[k]) e = hipMemcpyAsync(dst[k], tmp + 4 * k * nout, 16 * (size_t)nout, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(WT_E_HIP, std::string("wt_export_fields: ") + hipGetErrorString(e));
  return WT_OK;
}

// ---- probes, streamlines, particle tracers ------------------------------------------------------

int wt_probe(wt_solver* s, const float* pos4, int32_t count, float* out8) {
  if (!s || (count > 0 && (!pos4 || !out8))) return fail(WT_E_ARG, "null argument");
  if (count < 0) return fail(WT_E_ARG, "count must be >= 0");
  if (count == 0) return WT_OK;
  if (int r = wt::ensure_derived(s)) return r;
  float *dp = nullptr, *dout = nullptr;
  if (hipMalloc((void**)&dp, 16 * (size_t)count) != hipSuccess || hipMalloc((void**)&dout, 32 * (size_t)count) != hipSuccess) {
    if (dp) (void)hipFree(dp);
    return fail(WT_E_NOMEM, "probe buffers");
  }
  hipError_t e = hipMemcpyAsync(dp, pos4, 16 * (size_t)count, hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess) {
    wt::launch_probes(s->stream, s->vel, s->meanh, s->kind, s->dims.nx, s->dims.ny, s->dims.nz, dp, count, dout);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out8, dout, 32 * (size_t)count, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(dp);
  (void)hipFree(dout);
  if (e != hipSuccess) return fail(WT_E_HIP, std::string("wt_probe: ") + hipGetErrorString(e));
  return WT_OK;
}

int wt_streamlines(wt_solver* s, const wt_rake* rake, int32_t seeds, int32_t per_line, int32_t points, float h, int32_t use_mean,
                   float* out4) {
  if (!s || !rake || (seeds > 0 && points > 0 && !out4)) return fail(WT_E_ARG, "null argument");
  if (seeds < 0 || points < 0 || per_line < 0) return fail(WT_E_ARG, "seeds, per_line and points must be >= 0");
  if (!(h > 0.0f)) return fail(WT_E_ARG, "h must be > 0");
  if (seeds == 0 || points == 0) return WT_OK;
  if (use_mean) {
    if (int r = wt::ensure_derived(s)) return r;
  }
  const size_t bytes = 16 * (size_t)seeds * (size_t)points;
  float* dout = nullptr;
  if (hipMalloc((void**)&dout, bytes) != hipSuccess) return fail(WT_E_NOMEM, "streamline buffer");
  wt::launch_streamlines(s->stream, use_mean ? (const void*)s->meanh : (const void*)s->vel, s->kind, s->dims.nx, s->dims.ny,
                         s->dims.nz, *rake, seeds, per_line, points, h, (float)s->prm.u, dout);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out4, dout, bytes, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(dout);
  if (e != hipSuccess) return fail(WT_E_HIP, std::string("wt_streamlines: ") + hipGetErrorString(e));
  return WT_OK;
}

struct wt_particles {
  wt_solver* s = nullptr;
  wt_particle_opts o{};
  float* state = nullptr;  // count x float4
  float* hist = nullptr;   // trail x count x float4
  int head = 0;
  bool need_fill = true;
};

int wt_particles_create(wt_solver* s, const wt_particle_opts* o, wt_particles** out) {
  if (!s || !o || !out) return fail(WT_E_ARG, "null argument");
  if (o->count <= 0 || o->trail <= 0) return fail(WT_E_ARG, "count and trail must be > 0");
  if (o->emitter < 0 || o->emitter > 1) return fail(WT_E_ARG, "emitter must be 0 (inlet) or 1 (rake)");
  if ((int64_t)o->count * o->trail > ((int64_t)1 << 31)) return fail(WT_E_ARG, "count x trail too large");
  auto* p = new wt_particles();
  p->s = s;
  p->o = *o;
  if (hipMalloc((void**)&p->state, 16 * (size_t)o->count) != hipSuccess ||
      hipMalloc((void**)&p->hist, 16 * (size_t)o->count * (size_t)o->trail) != hipSuccess) {
    wt_particles_destroy(p);
    return fail(WT_E_NOMEM, "particle buffers");
  }
  // zeroed, as freshly created buffers are in the reference
  hipError_t e = hipMemsetAsync(p->state, 0, 16 * (size_t)o->count, s->stream);
  if (e == hipSuccess) e = hipMemsetAsync(p->hist, 0, 16 * (size_t)o->count * (size_t)o->trail, s->stream);
  if (e != hipSuccess) {
    wt_particles_destroy(p);
    return fail(WT_E_HIP, std::string("wt_particles_create: ") + hipGetErrorString(e));
  }
  *out = p;
  return WT_OK;
}

int wt_particles_destroy(wt_particles* p) {
  if (!p) return WT_OK;
  // hipFree waits for the device, so the solver (possibly destroyed already) is not touched here
  if (p->state) (void)hipFree(p->state);
  if (p->hist) (void)hipFree(p->hist);
  delete p;
  return WT_OK;
}

int wt_particles_update(wt_particles* p, float steps, uint32_t seed, int32_t fill, int32_t advance) {
  if (!p) return fail(WT_E_ARG, "null argument");
  const bool f = fill != 0 || p->need_fill;
  if (advance && !f) p->head = (p->head + 1) % p->o.trail;
  wt_solver* s = p->s;
  wt::launch_particles(s->stream, s->vel, s->kind, s->dims.nx, s->dims.ny, s->dims.nz, p->o, p->head, steps, seed, f ? 1 : 0,
                       advance ? 1 : 0, p->state, p->hist);
  HIP_TRY(hipGetLastError());
  p->need_fill = false;
  return WT_OK;
}

int wt_particles_read(wt_particles* p, float* state4, float* hist4, int32_t* head) {
  if (!p) return fail(WT_E_ARG, "null argument");
  hipStream_t st = p->s->stream;
  if (state4) HIP_TRY(hipMemcpyAsync(state4, p->state, 16 * (size_t)p->o.count, hipMemcpyDeviceToHost, st));
  if (hist4) HIP_TRY(hipMemcpyAsync(hist4, p->hist, 16 * (size_t)p->o.count * (size_t)p->o.trail, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (head) *head = p->head;
  return WT_OK;
}

int wt_profile_enable(wt_solver* s, int32_t on) {
  if (!s) return fail(WT_E_ARG, "null argument");
  HIP_TRY(hipStreamSynchronize(s->stream));
  if (on < WT_PROFILE_OFF || on > WT_PROFILE_SPAN) return fail(WT_E_ARG, "profile mode must be 0, 1 or 2");
  s->prof = on;
  s->ev_used = 0;
  s->ev_launches.clear();
  s->acc_bulk = s->acc_bnd = 0;
  s->acc_launches = 0;
  return WT_OK;
}

int wt_profile_read(wt_solver* s, double* bulk_ms, int64_t* launches, double* boundary_ms) {
  if (!s) return fail(WT_E_ARG, "null argument");
  HIP_TRY(hipStreamSynchronize(s->stream));
  for (size_t i = 0; i + 2 <= s->ev_used; i += 2) {
    float a = 0;
    HIP_TRY(hipEventElapsedTime(&a, s->ev[i], s->ev[i + 1]));
    s->acc_bulk += a;
    s->acc_launches += s->ev_launches[i / 2];
  }
  s->ev_used = 0;
  s->ev_launches.clear();
  if (bulk_ms) *bulk_ms = s->acc_bulk;
  if (launches) *launches = s->acc_launches;
  if (boundary_ms) *boundary_ms = s->acc_bnd;
  s->acc_bulk = s->acc_bnd = 0;
  s->acc_launches = 0;
  return WT_OK;
}

int wt_copy_probe(int64_t bytes, int32_t iters, double* gbs) {
  if (bytes < 16 || iters < 1 || !gbs) return fail(WT_E_ARG, "bad copy probe arguments");
  bytes &= ~(int64_t)15;
  void *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    if (a) (void)hipFree(a);
    (void)hipStreamDestroy(st);
    return fail(WT_E_NOMEM, "copy probe allocation failed");
  }
  HIP_TRY(hipMemsetAsync(a, 1, bytes, st));
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  launch_copy(st, a, b, bytes);  // warm-up
  HIP_TRY(hipEventRecord(e0, st));
  for (int k = 0; k < iters; k++) launch_copy(st, a, b, bytes);
  HIP_TRY(hipEventRecord(e1, st));
  HIP_TRY(hipEventSynchronize(e1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  *gbs = 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipStreamDestroy(st);
  return WT_OK;
}

int wt_set_tuning(wt_solver* s, int32_t key, int32_t value) {
  if (!s) return fail(WT_E_ARG, "null argument");
  if (key == WT_TUNE_ITEM_CHUNK) {
    if (value < 0) return fail(WT_E_ARG, "item chunk must be >= 0");
    s->item_chunk = value;
    return WT_OK;
  }
  if (key == WT_TUNE_CELLS_PER_THREAD) {
    if (value != 1 && value != 2) return fail(WT_E_ARG, "cells per thread must be 1 or 2");
    // FP16 / BF16 words hold two cells: those formats need pairs
    const int v = s->prec != WT_PREC_F32 ? 2 : value;
    if (v == s->cpt) return WT_OK;
    s->cpt = v;
    derive(s);
    return reclassify(s);  // the wall list is built per group of cells
  }
  return fail(WT_E_ARG, "unknown tuning key");
}

int wt_get_tuning(const wt_solver* s, int32_t key, int32_t* value) {
  if (!s || !value) return fail(WT_E_ARG, "null argument");
  if (key == WT_TUNE_ITEM_CHUNK) { *value = s->item_chunk; return WT_OK; }
  if (key == WT_TUNE_CELLS_PER_THREAD) { *value = s->P.cpt; return WT_OK; }
  return fail(WT_E_ARG, "unknown tuning key");
}

int wt_autotune(wt_solver* s, int32_t steps, wt_tune_result* out) {
  if (!s || !out) return fail(WT_E_ARG, "null argument");
  if (steps < 1) return fail(WT_E_ARG, "steps must be >= 1");
  if (s->pending_init) {
    if (int r = run_init(s)) return r;
  }
  std::memset(out, 0, sizeof(*out));
  // candidate runs of items per XCD: the plain order, one x-row, 8 rows, one z-plane, 4 planes,
  // one contiguous eighth per XCD
  const int cpt0 = s->P.cpt;
  const int64_t wpr = (s->dims.nx + 64 * cpt0 - 1) / (64 * cpt0);
  const int64_t cands[] = {4, wpr, 8 * wpr, wpr * s->dims.ny, 4 * wpr * s->dims.ny, 0};
  // the trials step the solver: keep its state and restore it afterwards
  const int64_t sbytes = s->prec == WT_PREC_F32 ? 4 : 2;
  const size_t fbytes = (size_t)((int64_t)s->q * s->n * sbytes);
  void* keep = nullptr;
  if (hipMalloc(&keep, fbytes) != hipSuccess) return fail(WT_E_NOMEM, "autotune snapshot allocation failed");
  const int parity0 = s->parity;
  const int64_t step0 = s->step;
  const int prof0 = s->prof;
  const int chunk0 = s->item_chunk;
  int rc = WT_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto run = [&]() -> int {
    HIP_TRY(hipMemcpyAsync(keep, s->f[parity0], fbytes, hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    s->prof = WT_PROFILE_OFF;
    double best = 1e30;
    int n = 0;
    for (int cpt : {2, 1}) {
      if (cpt == 2 && (s->dims.nx % 2 != 0)) continue;
      if (cpt == 1 && s->prec != WT_PREC_F32) continue;  // 16-bit words hold pairs
      if (int r = wt_set_tuning(s, WT_TUNE_CELLS_PER_THREAD, cpt)) return r;
      for (int64_t ck : cands) {
        if (n >= WT_TUNE_MAX_TRIALS) break;
        s->item_chunk = (int)ck;
        if (int r = do_steps(s, 2)) return r;  // warm-up
        HIP_TRY(hipEventRecord(e0, s->stream));
        if (int r = do_steps(s, steps)) return r;
        HIP_TRY(hipEventRecord(e1, s->stream));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        const double per = ms / steps;
        out->trial_chunk[n] = (int32_t)ck;
        out->trial_cpt[n] = cpt;
        out->trial_ms[n] = per;
        n++;
        if (per < best) {
          best = per;
          out->chunk = (int32_t)ck;
          out->cells_per_thread = cpt;
          out->ms_per_step = per;
        }
      }
    }
    out->trials = n;
    return WT_OK;
  };
  rc = run();
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc == WT_OK) rc = wt_set_tuning(s, WT_TUNE_CELLS_PER_THREAD, out->cells_per_thread);
  if (rc == WT_OK) {
    s->item_chunk = out->chunk;
  } else {
    // a failed trial leaves the handle's launch configuration (and its wall list) as it was
    const std::string msg = g_err;
    s->item_chunk = chunk0;
    (void)wt_set_tuning(s, WT_TUNE_CELLS_PER_THREAD, cpt0);
    g_err = msg;
  }
  // restore the state the trials stepped away from
  const hipError_t e = hipMemcpyAsync(s->f[parity0], keep, fbytes, hipMemcpyDeviceToDevice, s->stream);
  const hipError_t e2 = hipStreamSynchronize(s->stream);
  (void)hipFree(keep);
  s->parity = parity0;
  s->step = step0;
  s->prof = prof0;
  if (rc != WT_OK) return rc;
  if (e != hipSuccess || e2 != hipSuccess) return fail(WT_E_HIP, "autotune state restore failed");
  return WT_OK;
}

int wt_work_split(const wt_solver* s, int64_t* bulk_cells, int64_t* boundary_cells) {
  if (!s) return fail(WT_E_ARG, "null argument");
  if (bulk_cells) *bulk_cells = s->nfast;
  if (boundary_cells) *boundary_cells = s->nwall;
  return WT_OK;
}

// ---- voxelizers ----------------------------------------------------------------------------

int wt_voxelize_sdf(wt_solver* s, const wt_prim* prims, int32_t count, wt_voxel_info* info) {
  if (!s || (count > 0 && !prims)) return fail(WT_E_ARG, "null argument");
  if (count < 0 || count > WT_MAX_PRIMS) return fail(WT_E_ARG, "primitive count must be in [0, 64]");
  VoxCtx v{s->stream, s->dims.nx, s->dims.ny, s->dims.nz, s->n, s->flags, s->sdf, &s->W};
  int64_t h[8];
  if (int r = voxelize_sdf(v, prims, count, h)) return fail(WT_E_HIP, "analytic voxelizer failed");
  if (int r = reclassify(s)) return r;
  if (info) {
    info->frontal = h[0];
    info->solid_cells = h[1];
    for (int a = 0; a < 3; a++) {
      info->min[a] = h[1] > 0 ? (int32_t)h[2 + a] : 0;
      info->max[a] = h[1] > 0 ? (int32_t)h[5 + a] + 1 : 0;
    }
  }
  return WT_OK;
}

int wt_voxelize_mesh(wt_solver* s, const float* positions, int32_t n_vertices, const uint32_t* indices, int32_t n_triangles,
                     const float* matrix16, const float* wheels, int32_t n_wheels, wt_voxel_info* info) {
  if (!s || !matrix16 || (n_vertices > 0 && !positions) || (n_triangles > 0 && !indices)) return fail(WT_E_ARG, "null argument");
  if (n_vertices < 0 || n_triangles < 0) return fail(WT_E_ARG, "negative size");
  for (int64_t i = 0; i < 3 * (int64_t)n_triangles; i++)
    if (indices[i] >= (uint32_t)n_vertices) return fail(WT_E_ARG, "triangle index out of range");
  // VoxelizerGPU.ts:49: the call always replaces the solver's wheel list (an empty list clears it)
  if (int r = wt_set_wheels(s, wheels, wheels ? n_wheels : 0)) return r;
  VoxCtx v{s->stream, s->dims.nx, s->dims.ny, s->dims.nz, s->n, s->flags, s->sdf, &s->W};
  int64_t h[8];
  if (int r = voxelize_mesh(v, positions, n_vertices, indices, n_triangles, matrix16, h)) return fail(WT_E_HIP, g_err.empty() ? "mesh voxelizer failed" : g_err);
  if (int r = reclassify(s)) return r;
  if (info) {
    info->frontal = h[0];
    info->solid_cells = h[1];
    for (int a = 0; a < 3; a++) {
      info->min[a] = h[1] > 0 ? (int32_t)h[2 + a] : 0;
      info->max[a] = h[1] > 0 ? (int32_t)h[5 + a] + 1 : 0;
    }
  }
  return WT_OK;
}

// ---- batched sweeps ----------------------------------------------------------------------

int wt_batch_create(wt_solver* const* members, int32_t count, wt_batch** out) {
  if (!members || !out || count < 1) return fail(WT_E_ARG, "bad batch member list");
  *out = nullptr;
  const wt_solver* m0 = members[0];
  if (!m0) return fail(WT_E_ARG, "null batch member");
  for (int32_t k = 0; k < count; k++) {
    const wt_solver* s = members[k];
    if (!s) return fail(WT_E_ARG, "null batch member");
    if (s->q != m0->q || s->prec != m0->prec || cells_per_thread(s->P) != cells_per_thread(m0->P) ||
        (s->P.collision != 1) != (m0->P.collision != 1))
      return fail(WT_E_ARG, "batch members must share lattice, storage precision, cells per thread and collision");
    if (s->P.closure == WT_CLOSURE_MLP && s->mlp)
      return fail(WT_E_STATE, "a batch member uses the learned closure, which the batched step does not run");
    if (s->parity != m0->parity) return fail(WT_E_STATE, "batch members must have been stepped equally often (same parity)");
    if (s->device != m0->device) return fail(WT_E_ARG, "batch members must live on one device");
    for (int32_t j = 0; j < k; j++)
      if (members[j] == s) return fail(WT_E_ARG, "a solver appears twice in the batch");
  }
  auto* b = new wt_batch();
  b->m.assign(members, members + count);
  b->q = m0->q;
  b->prec = m0->prec;
  b->cpt = cells_per_thread(m0->P);
  b->reg = m0->P.collision != 1;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess) {
    wt_batch_destroy(b);
    return fail(WT_E_HIP, "batch stream creation failed");
  }
  *out = b;
  return WT_OK;
}

int wt_batch_destroy(wt_batch* b) {
  if (!b) return WT_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_drop_graph(b);
  if (b->d_desc) (void)hipFree(b->d_desc);
  if (b->d_map) (void)hipFree(b->d_map);
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return WT_OK;
}

int wt_batch_set_chunk(wt_batch* b, int32_t chunk_steps, int32_t use_graph) {
  if (!b || chunk_steps < 2) return fail(WT_E_ARG, "chunk must be >= 2 steps");
  if (b->stream) HIP_TRY(hipStreamSynchronize(b->stream));
  batch_drop_graph(b);
  b->chunk = (chunk_steps + 1) / 2 * 2;  // even: a replay leaves the parity unchanged
  b->use_graph = use_graph != 0;
  return WT_OK;
}

int wt_batch_run(wt_batch* b, int32_t steps) {
  if (!b) return fail(WT_E_ARG, "null argument");
  if (steps < 0) return fail(WT_E_ARG, "steps must be >= 0");
  const int parity = b->m[0]->parity;
  for (wt_solver* s : b->m) {
    if (s->parity != parity) return fail(WT_E_STATE, "batch members were stepped apart (parity differs)");
    if (s->pending_init) {
      if (int r = run_init(s)) return r;
    }
    HIP_TRY(hipStreamSynchronize(s->stream));  // uploads, preparation and init are done
  }
  if (int r = batch_prepare(b)) return r;
  int done = 0;
  if (b->use_graph && steps >= b->chunk) {
    if (!b->gexec[parity]) {
      // one captured chunk of `chunk` launches, parities alternating from the current one; chunk
      // is even, so every replay starts from the same parity
      HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      for (int k = 0; k < b->chunk; k++) batch_launch(b, (parity + k) & 1);
      hipGraph_t g = nullptr;
      const hipError_t ec = hipStreamEndCapture(b->stream, &g);
      if (ec != hipSuccess) return fail(WT_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
      b->graph[parity] = g;
      HIP_TRY(hipGraphInstantiate(&b->gexec[parity], g, nullptr, nullptr, 0));
    }
    for (; done + b->chunk <= steps; done += b->chunk) HIP_TRY(hipGraphLaunch(b->gexec[parity], b->stream));
  }
  for (; done < steps; done++) batch_launch(b, (parity + done) & 1);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(b->done, b->stream));
  for (wt_solver* s : b->m) {
    HIP_TRY(hipStreamWaitEvent(s->stream, b->done, 0));  // later work on a member follows the batch
    s->parity = (s->parity + steps) & 1;
    s->step += steps;
  }
  return WT_OK;
}

int wt_batch_sync(wt_batch* b) {
  if (!b) return fail(WT_E_ARG, "null argument");
  HIP_TRY(hipStreamSynchronize(b->stream));
  return WT_OK;
}

int wt_batch_stats(wt_batch* b, wt_stats* out) {
  if (!b || !out) return fail(WT_E_ARG, "null argument");
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (size_t k = 0; k < b->m.size(); k++) {
    if (int r = do_sample(b->m[k])) return r;
    decode_stats(b->m[k], &out[k]);
  }
  return WT_OK;
}

}  // extern "C"
