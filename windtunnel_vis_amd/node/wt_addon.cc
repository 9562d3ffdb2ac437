// N-API addon binding libwt.so for the Node host (solver.mjs). Plain C N-API (node_api.h), no
// extra packages. Each solver handle is a JS external; stepAndSample runs as async work off the
// JS thread and resolves a Promise, like the reference's mapAsync-based stepAndSample
// (src/solver/SolverGPU.ts:292-309).
#include <node_api.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "wt.h"

namespace {

#define NAPI_CALL(env, call)                                              \
  do {                                                                    \
    if ((call) != napi_ok) {                                              \
      napi_throw_error(env, nullptr, "N-API call failed: " #call);        \
      return nullptr;                                                     \
    }                                                                     \
  } while (0)

napi_value throw_wt(napi_env env, int rc) {
  std::string msg = std::string("libwt error ") + std::to_string(rc) + ": " + wt_last_error();
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

// One solver handle. libwt handles are not re-entrant, so every call on one runs on the JS thread,
// in call order: stepAndSample enqueues its steps, reduction and read there
// (wt_step_and_sample_async) and leaves only the wait for the read to a worker thread
// (wt_sample_wait, which touches nothing but the read's slot). A destroy() while samples are in
// flight only marks the handle (`closing`); the last sample's completion destroys it. Each
// in-flight sample holds a reference to the JS handle, so the finalizer cannot run under it.
struct Handle {
  wt_solver* s = nullptr;
  std::mutex mu;         // held by the JS-thread calls and by the completion that destroys
  int pending = 0;       // samples in flight (changed on the JS thread only)
  bool closing = false;  // destroy() called while samples were in flight
};

void finalize_handle(napi_env, void* data, void*) {
  auto* h = static_cast<Handle*>(data);
  if (h->s) wt_destroy(h->s);
  delete h;
}

Handle* get_handle(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "expected a solver handle");
    return nullptr;
  }
  auto* h = static_cast<Handle*>(p);
  if (!h->s || h->closing) {
    napi_throw_error(env, nullptr, "solver was destroyed");
    return nullptr;
  }
  return h;
}

// A solver handle with its lock held for the scope (see Handle).
struct HandleLock {
  Handle* h = nullptr;
  std::unique_lock<std::mutex> lk;
  HandleLock(napi_env env, napi_value v) : h(get_handle(env, v)) {
    if (h) lk = std::unique_lock<std::mutex>(h->mu);
  }
  explicit operator bool() const { return h != nullptr; }
};

bool get_prop_num(napi_env env, napi_value obj, const char* key, double* out) {
  bool has = false;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return false;
  napi_value v;
  napi_get_named_property(env, obj, key, &v);
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_boolean) {
    bool b;
    napi_get_value_bool(env, v, &b);
    *out = b ? 1.0 : 0.0;
    return true;
  }
  if (t != napi_number) return false;
  napi_get_value_double(env, v, out);
  return true;
}

// params object: numeric fields with the C names; enums already mapped to integers by solver.mjs.
uint32_t parse_params(napi_env env, napi_value obj, wt_params* p) {
  uint32_t mask = 0;
  double v;
  struct F { const char* k; uint32_t bit; int kind; void* dst; };
  F fields[] = {
      {"u", WT_P_U, 0, &p->u}, {"nu", WT_P_NU, 0, &p->nu}, {"cs", WT_P_CS, 0, &p->cs},
      {"ground", WT_P_GROUND, 1, &p->ground}, {"sides", WT_P_SIDES, 1, &p->sides},
      {"spongeNu", WT_P_SPONGE_NU, 0, &p->sponge_nu}, {"spongeStart", WT_P_SPONGE_START, 0, &p->sponge_start},
      {"spongeIn", WT_P_SPONGE_IN, 0, &p->sponge_in}, {"tauWall", WT_P_TAU_WALL, 0, &p->tau_wall},
      {"interp", WT_P_INTERP, 1, &p->interp}, {"collision", WT_P_COLLISION, 1, &p->collision},
      {"closure", WT_P_CLOSURE, 1, &p->closure}, {"emaAlpha", WT_P_EMA_ALPHA, 0, &p->ema_alpha},
  };
  for (auto& f : fields) {
    if (!get_prop_num(env, obj, f.k, &v)) continue;
    if (f.kind == 0) *(double*)f.dst = v;
    else *(int32_t*)f.dst = (int32_t)v;
    mask |= f.bit;
  }
  if (get_prop_num(env, obj, "precision", &v)) p->precision = (int32_t)v;
  return mask;
}

template <class T>
bool typed_array(napi_env env, napi_value v, napi_typedarray_type want, T** data, size_t* len) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) return false;
  napi_typedarray_type t;
  void* d;
  napi_value ab;
  size_t off;
  napi_get_typedarray_info(env, v, &t, len, &d, &ab, &off);
  if (t != want) return false;
  *data = static_cast<T*>(d);
  return true;
}

napi_value make_typed(napi_env env, napi_typedarray_type t, size_t len, size_t elem, void** data) {
  napi_value ab, out;
  if (napi_create_arraybuffer(env, len * elem, data, &ab) != napi_ok) return nullptr;
  if (napi_create_typedarray(env, t, len, ab, 0, &out) != napi_ok) return nullptr;
  return out;
}

void set_num(napi_env env, napi_value obj, const char* k, double v) {
  napi_value x;
  napi_create_double(env, v, &x);
  napi_set_named_property(env, obj, k, x);
}

napi_value stats_object(napi_env env, const wt_stats& st) {
  napi_value o, b;
  napi_create_object(env, &o);
  set_num(env, o, "fx", st.fx);
  set_num(env, o, "fy", st.fy);
  set_num(env, o, "fz", st.fz);
  set_num(env, o, "maxU", st.max_u);
  set_num(env, o, "rhoRef", st.rho_ref);
  napi_get_boolean(env, st.unstable != 0, &b);
  napi_set_named_property(env, o, "unstable", b);
  set_num(env, o, "step", (double)st.step);
  return o;
}

napi_value voxel_object(napi_env env, const wt_voxel_info& vi) {
  napi_value o, mn, mx;
  napi_create_object(env, &o);
  set_num(env, o, "frontal", (double)vi.frontal);
  set_num(env, o, "solidCells", (double)vi.solid_cells);
  napi_create_array_with_length(env, 3, &mn);
  napi_create_array_with_length(env, 3, &mx);
  for (uint32_t a = 0; a < 3; a++) {
    napi_value x;
    napi_create_int32(env, vi.min[a], &x);
    napi_set_element(env, mn, a, x);
    napi_create_int32(env, vi.max[a], &x);
    napi_set_element(env, mx, a, x);
  }
  napi_set_named_property(env, o, "min", mn);
  napi_set_named_property(env, o, "max", mx);
  return o;
}

napi_value args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr);
  if (argc < want) {
    napi_throw_type_error(env, nullptr, "not enough arguments");
    return nullptr;
  }
  return argv[0];
}

// create(dims {nx, ny, nz}, params) -> handle
napi_value Create(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  double nx = 0, ny = 0, nz = 0;
  get_prop_num(env, argv[0], "nx", &nx);
  get_prop_num(env, argv[0], "ny", &ny);
  get_prop_num(env, argv[0], "nz", &nz);
  wt_dims d{(int32_t)nx, (int32_t)ny, (int32_t)nz};
  wt_params p{};
  p.tau_wall = 0.5;
  p.interp = 1;
  parse_params(env, argv[1], &p);
  wt_solver* s = nullptr;
  int rc = wt_create(&d, &p, &s);
  if (rc) return throw_wt(env, rc);
  auto* h = new Handle{s};
  napi_value ext;
  NAPI_CALL(env, napi_create_external(env, h, finalize_handle, nullptr, &ext));
  return ext;
}

napi_value Destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  void* p = nullptr;
  napi_get_value_external(env, argv[0], &p);
  auto* h = static_cast<Handle*>(p);
  if (h && h->s && !h->closing) {
    if (h->pending > 0) {
      h->closing = true;  // the last in-flight sample destroys it
    } else {
      std::lock_guard<std::mutex> g(h->mu);
      wt_destroy(h->s);
      h->s = nullptr;
    }
  }
  return nullptr;
}

napi_value SetParams(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  Handle* h = lk.h;
  wt_params p{};
  wt_get_params(h->s, &p);
  const uint32_t mask = parse_params(env, argv[1], &p);
  int rc = wt_set_params(h->s, &p, mask);
  if (rc) return throw_wt(env, rc);
  return nullptr;
}

napi_value Simple(napi_env env, napi_callback_info info, int (*fn)(wt_solver*)) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  Handle* h = lk.h;
  int rc = fn(h->s);
  if (rc) return throw_wt(env, rc);
  return nullptr;
}
napi_value Reset(napi_env env, napi_callback_info info) { return Simple(env, info, wt_reset); }
napi_value ClearStats(napi_env env, napi_callback_info info) { return Simple(env, info, wt_clear_stats); }
napi_value Sync(napi_env env, napi_callback_info info) { return Simple(env, info, wt_sync); }

napi_value Info(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  Handle* h = lk.h;
  wt_info i{};
  int rc = wt_info_get(h->s, &i);
  if (rc) return throw_wt(env, rc);
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "nx", i.dims.nx);
  set_num(env, o, "ny", i.dims.ny);
  set_num(env, o, "nz", i.dims.nz);
  set_num(env, o, "n", (double)i.n);
  set_num(env, o, "q", i.q);
  set_num(env, o, "precision", i.precision);
  set_num(env, o, "step", (double)i.step);
  set_num(env, o, "wallCells", (double)i.wall_cells);
  set_num(env, o, "deviceBytes", (double)i.device_bytes);
  set_num(env, o, "tau0", wt_tau0(h->s));
  return o;
}

napi_value UploadFlags(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  Handle* h = lk.h;
  uint32_t* d;
  size_t len;
  if (!typed_array(env, argv[1], napi_uint32_array, &d, &len)) {
    napi_throw_type_error(env, nullptr, "flags must be a Uint32Array");
    return nullptr;
  }
  wt_info i{};
  wt_info_get(h->s, &i);
  if ((int64_t)len != i.n) {
    napi_throw_range_error(env, nullptr, "flags length must equal the cell count");
    return nullptr;
  }
  int rc = wt_upload_flags(h->s, d);
  if (rc) return throw_wt(env, rc);
  return nullptr;
}

napi_value ReadFlags(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  Handle* h = lk.h;
  wt_info i{};
  wt_info_get(h->s, &i);
  void* data;
  napi_value out = make_typed(env, napi_uint32_array, (size_t)i.n, 4, &data);
  int rc = wt_read_flags(h->s, (uint32_t*)data);
  if (rc) return throw_wt(env, rc);
  return out;
}

napi_value UploadSdf(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  Handle* h = lk.h;
  float* d;
  size_t len;
  if (!typed_array(env, argv[1], napi_float32_array, &d, &len)) {
    napi_throw_type_error(env, nullptr, "sdf must be a Float32Array");
    return nullptr;
  }
  wt_info i{};
  wt_info_get(h->s, &i);
  if ((int64_t)len != i.n) {
    napi_throw_range_error(env, nullptr, "sdf length must equal the cell count");
    return nullptr;
  }
  int rc = wt_upload_sdf(h->s, d);
  if (rc) return throw_wt(env, rc);
  return nullptr;
}

// setWheels(h, Float32Array of 8 per wheel)
napi_value SetWheels(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  Handle* h = lk.h;
  float* d;
  size_t len;
  if (!typed_array(env, argv[1], napi_float32_array, &d, &len) || len % 8 != 0) {
    napi_throw_type_error(env, nullptr, "wheels must be a Float32Array of 8 floats per wheel");
    return nullptr;
  }
  int rc = wt_set_wheels(h->s, len ? d : nullptr, (int32_t)(len / 8));
  if (rc) return throw_wt(env, rc);
  return nullptr;
}

napi_value Step(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  Handle* h = lk.h;
  int32_t n = 0;
  napi_get_value_int32(env, argv[1], &n);
  int rc = wt_step(h->s, n);
  if (rc) return throw_wt(env, rc);
  return nullptr;
}

struct SampleWork {
  napi_async_work work;
  napi_deferred deferred;
  Handle* h;
  napi_ref ref;  // keeps the JS handle (and with it the Handle) alive while the work is in flight
  wt_solver* s;  // the solver at post time (freed only after the last completion)
  int64_t ticket;
  int rc;
  std::string err;
  wt_stats st;
};

void sample_execute(napi_env, void* data) {
  auto* w = static_cast<SampleWork*>(data);
  w->rc = wt_sample_wait(w->s, w->ticket, &w->st);
  if (w->rc) w->err = wt_last_error();
}

void sample_complete(napi_env env, napi_status, void* data) {
  auto* w = static_cast<SampleWork*>(data);
  Handle* h = w->h;
  if (--h->pending == 0 && h->closing && h->s) {
    std::lock_guard<std::mutex> g(h->mu);
    wt_destroy(h->s);
    h->s = nullptr;
  }
  napi_delete_reference(env, w->ref);
  if (w->rc) {
    napi_value msg, err;
    std::string m = "libwt error " + std::to_string(w->rc) + ": " + w->err;
    napi_create_string_utf8(env, m.c_str(), NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, nullptr, msg, &err);
    napi_reject_deferred(env, w->deferred, err);
  } else {
    napi_resolve_deferred(env, w->deferred, stats_object(env, w->st));
  }
  napi_delete_async_work(env, w->work);
  delete w;
}

// stepAndSample(h, steps) -> Promise<SolverStats>
napi_value StepAndSample(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  Handle* h = lk.h;
  int32_t steps = 0;
  napi_get_value_int32(env, argv[1], &steps);
  int64_t ticket = 0;
  int rc = wt_step_and_sample_async(h->s, steps, &ticket);  // enqueued now, in call order
  if (rc) return throw_wt(env, rc);
  auto* w = new SampleWork();
  w->h = h;
  w->s = h->s;
  w->ticket = ticket;
  napi_value promise, name;
  NAPI_CALL(env, napi_create_promise(env, &w->deferred, &promise));
  NAPI_CALL(env, napi_create_reference(env, argv[0], 1, &w->ref));
  napi_create_string_utf8(env, "wt_step_and_sample", NAPI_AUTO_LENGTH, &name);
  NAPI_CALL(env, napi_create_async_work(env, nullptr, name, sample_execute, sample_complete, w, &w->work));
  h->pending++;
  NAPI_CALL(env, napi_queue_async_work(env, w->work));
  return promise;
}

// stepAndSampleSync(h, steps) -> SolverStats (blocking variant)
napi_value StepAndSampleSync(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  Handle* h = lk.h;
  int32_t n = 0;
  napi_get_value_int32(env, argv[1], &n);
  wt_stats st{};
  int rc = wt_step_and_sample(h->s, n, &st);
  if (rc) return throw_wt(env, rc);
  return stats_object(env, st);
}

napi_value ReadPopulations(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  Handle* h = lk.h;
  wt_info i{};
  wt_info_get(h->s, &i);
  void* data;
  napi_value out = make_typed(env, napi_float32_array, (size_t)(i.q * i.n), 4, &data);
  int rc = wt_read_populations(h->s, (float*)data);
  if (rc) return throw_wt(env, rc);
  return out;
}

napi_value UploadPopulations(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  Handle* h = lk.h;
  float* d;
  size_t len;
  wt_info i{};
  wt_info_get(h->s, &i);
  if (!typed_array(env, argv[1], napi_float32_array, &d, &len) || (int64_t)len != i.q * i.n) {
    napi_throw_type_error(env, nullptr, "populations must be a Float32Array of q * n values");
    return nullptr;
  }
  int rc = wt_upload_populations(h->s, d);
  if (rc) return throw_wt(env, rc);
  return nullptr;
}

// voxelizeSdf(h, Float32Array prims packed as 20 floats each: kind, flags, c[3], rot[9], half[3], radius, param, pad)
napi_value VoxelizeSdf(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  Handle* h = lk.h;
  float* d;
  size_t len;
  if (!typed_array(env, argv[1], napi_float32_array, &d, &len) || len % 20 != 0) {
    napi_throw_type_error(env, nullptr, "prims must be a Float32Array of 20 floats per primitive");
    return nullptr;
  }
  std::vector<wt_prim> prims(len / 20);
  for (size_t k = 0; k < prims.size(); k++) {
    const float* e = d + 20 * k;
    wt_prim& p = prims[k];
    p.kind = (int32_t)e[0];
    p.flags = (int32_t)e[1];
    for (int a = 0; a < 3; a++) p.center[a] = e[2 + a];
    for (int a = 0; a < 9; a++) p.rot[a] = e[5 + a];
    for (int a = 0; a < 3; a++) p.half[a] = e[14 + a];
    p.radius = e[17];
    p.param = e[18];
  }
  wt_voxel_info vi{};
  int rc = wt_voxelize_sdf(h->s, prims.data(), (int32_t)prims.size(), &vi);
  if (rc) return throw_wt(env, rc);
  return voxel_object(env, vi);
}

// voxelizeMesh(h, positions Float32Array, indices Uint32Array, matrix Float32Array(16), wheels Float32Array)
napi_value VoxelizeMesh(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  if (!args(env, info, 5, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  Handle* h = lk.h;
  float *pos, *m, *wh;
  uint32_t* idx;
  size_t np, ni, nm, nw;
  if (!typed_array(env, argv[1], napi_float32_array, &pos, &np) || !typed_array(env, argv[2], napi_uint32_array, &idx, &ni) ||
      !typed_array(env, argv[3], napi_float32_array, &m, &nm) || nm != 16 || !typed_array(env, argv[4], napi_float32_array, &wh, &nw) ||
      np % 3 || ni % 3 || nw % 8) {
    napi_throw_type_error(env, nullptr, "voxelizeMesh(h, Float32Array pos, Uint32Array idx, Float32Array(16), Float32Array wheels)");
    return nullptr;
  }
  wt_voxel_info vi{};
  int rc = wt_voxelize_mesh(h->s, pos, (int32_t)(np / 3), idx, (int32_t)(ni / 3), m, nw ? wh : nullptr, (int32_t)(nw / 8), &vi);
  if (rc) return throw_wt(env, rc);
  return voxel_object(env, vi);
}

// derive(h): derived pass over the last sample's fields (vorticity, Q, TI, mean copy)
napi_value Derive(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  Handle* h = lk.h;
  int rc = wt_derive(h->s);
  if (rc) return throw_wt(env, rc);
  return nullptr;
}

// exportFields(h, stride) -> {dims: [ox, oy, oz], vel, mean, vort} (Float32Array, 4 per point)
napi_value ExportFields(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  Handle* h = lk.h;
  int32_t stride = 0;
  napi_get_value_int32(env, argv[1], &stride);
  int32_t o[3];
  int rc = wt_export_dims(h->s, stride, o);
  if (rc) return throw_wt(env, rc);
  const size_t m = (size_t)o[0] * o[1] * o[2];
  void *dv, *dm, *dw;
  napi_value vel = make_typed(env, napi_float32_array, 4 * m, 4, &dv);
  napi_value mean = make_typed(env, napi_float32_array, 4 * m, 4, &dm);
  napi_value vort = make_typed(env, napi_float32_array, 4 * m, 4, &dw);
  if (!vel || !mean || !vort) {
    napi_throw_error(env, nullptr, "exportFields: out of memory");
    return nullptr;
  }
  rc = wt_export_fields(h->s, stride, (float*)dv, (float*)dm, (float*)dw, o);
  if (rc) return throw_wt(env, rc);
  napi_value out, dims;
  napi_create_object(env, &out);
  napi_create_array_with_length(env, 3, &dims);
  for (uint32_t k = 0; k < 3; k++) {
    napi_value x;
    napi_create_int32(env, o[k], &x);
    napi_set_element(env, dims, k, x);
  }
  napi_set_named_property(env, out, "dims", dims);
  napi_set_named_property(env, out, "vel", vel);
  napi_set_named_property(env, out, "mean", mean);
  napi_set_named_property(env, out, "vort", vort);
  return out;
}

// probe(h, Float32Array pos4) -> Float32Array (8 per point: field ux, uy, uz, rho, then the mean)
napi_value Probe(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  Handle* h = lk.h;
  float* pos;
  size_t np;
  if (!typed_array(env, argv[1], napi_float32_array, &pos, &np) || np % 4) {
    napi_throw_type_error(env, nullptr, "probe(h, Float32Array of 4 floats per point)");
    return nullptr;
  }
  void* d;
  napi_value out = make_typed(env, napi_float32_array, 2 * np, 4, &d);
  if (!out) {
    napi_throw_error(env, nullptr, "out of memory");
    return nullptr;
  }
  int rc = wt_probe(h->s, pos, (int32_t)(np / 4), (float*)d);
  if (rc) return throw_wt(env, rc);
  return out;
}

bool get_rake(napi_env env, napi_value v, wt_rake* r) {
  float* e;
  size_t n;
  if (!typed_array(env, v, napi_float32_array, &e, &n) || n != 12) return false;
  for (int k = 0; k < 4; k++) {
    r->a[k] = e[k];
    r->b[k] = e[4 + k];
    r->d[k] = e[8 + k];
  }
  return true;
}

// streamlines(h, rake Float32Array(12) = a, b, d (xyzw each), seeds, perLine, points, step, useMean)
//   -> Float32Array (seeds x points x 4: x, y, z, speed or -1)
napi_value Streamlines(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  if (!args(env, info, 7, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  Handle* h = lk.h;
  wt_rake r{};
  if (!get_rake(env, argv[1], &r)) {
    napi_throw_type_error(env, nullptr, "streamlines: rake must be a Float32Array(12): a, b, d as xyzw");
    return nullptr;
  }
  int32_t seeds = 0, per = 0, points = 0;
  double step = 0.5;
  bool mean = false;
  napi_get_value_int32(env, argv[2], &seeds);
  napi_get_value_int32(env, argv[3], &per);
  napi_get_value_int32(env, argv[4], &points);
  napi_get_value_double(env, argv[5], &step);
  napi_get_value_bool(env, argv[6], &mean);
  if (seeds < 1 || points < 1) {
    napi_throw_range_error(env, nullptr, "streamlines: seeds and points must be >= 1");
    return nullptr;
  }
  void* d;
  napi_value out = make_typed(env, napi_float32_array, (size_t)seeds * points * 4, 4, &d);
  if (!out) {
    napi_throw_error(env, nullptr, "out of memory");
    return nullptr;
  }
  int rc = wt_streamlines(h->s, &r, seeds, per > 0 ? per : seeds, points, (float)step, mean ? 1 : 0, (float*)d);
  if (rc) return throw_wt(env, rc);
  return out;
}

// A particle system keeps its solver's external alive (a reference) and must be destroyed before
// the solver is; after the solver's own destroy() it refuses to update or read.
struct PHandle {
  wt_particles* p = nullptr;
  Handle* owner = nullptr;
  napi_ref owner_ref = nullptr;
  int32_t count = 0, trail = 0;
};

void finalize_particles(napi_env env, void* data, void*) {
  auto* ph = static_cast<PHandle*>(data);
  if (ph->p) wt_particles_destroy(ph->p);  // touches no solver memory
  if (ph->owner_ref) napi_delete_reference(env, ph->owner_ref);
  delete ph;
}

PHandle* get_particles(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "expected a particle handle");
    return nullptr;
  }
  auto* ph = static_cast<PHandle*>(p);
  if (!ph->p) {
    napi_throw_error(env, nullptr, "particle system was destroyed");
    return nullptr;
  }
  if (!ph->owner || !ph->owner->s || ph->owner->closing) {
    napi_throw_error(env, nullptr, "the particle system's solver was destroyed");
    return nullptr;
  }
  return ph;
}

// particlesCreate(h, {count, trail, emitter: 0 inlet | 1 rake, nozzles, lines, maxAge, rake: Float32Array(12)})
napi_value ParticlesCreate(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  Handle* h = lk.h;
  wt_particle_opts o{};
  double v = 0;
  o.count = get_prop_num(env, argv[1], "count", &v) ? (int32_t)v : 0;
  o.trail = get_prop_num(env, argv[1], "trail", &v) ? (int32_t)v : 16;
  o.emitter = get_prop_num(env, argv[1], "emitter", &v) ? (int32_t)v : 0;
  o.nozzles = get_prop_num(env, argv[1], "nozzles", &v) ? (int32_t)v : 16;
  o.lines = get_prop_num(env, argv[1], "lines", &v) ? (int32_t)v : 1;
  o.max_age = get_prop_num(env, argv[1], "maxAge", &v) ? (float)v : 400.0f;
  bool has = false;
  napi_has_named_property(env, argv[1], "rake", &has);
  if (has) {
    napi_value rv;
    napi_get_named_property(env, argv[1], "rake", &rv);
    if (!get_rake(env, rv, &o.rake)) {
      napi_throw_type_error(env, nullptr, "particlesCreate: rake must be a Float32Array(12)");
      return nullptr;
    }
  }
  wt_particles* p = nullptr;
  int rc = wt_particles_create(h->s, &o, &p);
  if (rc) return throw_wt(env, rc);
  auto* ph = new PHandle{p, h, nullptr, o.count, o.trail};
  napi_create_reference(env, argv[0], 1, &ph->owner_ref);
  napi_value ext;
  NAPI_CALL(env, napi_create_external(env, ph, finalize_particles, nullptr, &ext));
  return ext;
}

napi_value ParticlesDestroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  void* p = nullptr;
  napi_get_value_external(env, argv[0], &p);
  auto* ph = static_cast<PHandle*>(p);
  if (ph && ph->p) {
    wt_particles_destroy(ph->p);
    ph->p = nullptr;
    if (ph->owner_ref) {
      napi_delete_reference(env, ph->owner_ref);
      ph->owner_ref = nullptr;
    }
  }
  return nullptr;
}

// particlesUpdate(ph, steps, seed, fill, advance)
napi_value ParticlesUpdate(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  if (!args(env, info, 5, argv)) return nullptr;
  PHandle* ph = get_particles(env, argv[0]);
  if (!ph) return nullptr;
  std::lock_guard<std::mutex> g(ph->owner->mu);  // the tracers run on the solver's stream and fields
  double steps = 0;
  uint32_t seed = 0;
  bool fill = false, advance = true;
  napi_get_value_double(env, argv[1], &steps);
  napi_get_value_uint32(env, argv[2], &seed);
  napi_get_value_bool(env, argv[3], &fill);
  napi_get_value_bool(env, argv[4], &advance);
  int rc = wt_particles_update(ph->p, (float)steps, seed, fill ? 1 : 0, advance ? 1 : 0);
  if (rc) return throw_wt(env, rc);
  return nullptr;
}

// particlesRead(ph) -> {state: Float32Array(count x 4), trail: Float32Array(trail x count x 4), head}
napi_value ParticlesRead(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  PHandle* ph = get_particles(env, argv[0]);
  if (!ph) return nullptr;
  std::lock_guard<std::mutex> g(ph->owner->mu);  // the tracers run on the solver's stream and fields
  void *ds, *dh;
  napi_value st = make_typed(env, napi_float32_array, (size_t)ph->count * 4, 4, &ds);
  napi_value tr = make_typed(env, napi_float32_array, (size_t)ph->trail * ph->count * 4, 4, &dh);
  if (!st || !tr) {
    napi_throw_error(env, nullptr, "out of memory");
    return nullptr;
  }
  int32_t head = 0;
  int rc = wt_particles_read(ph->p, (float*)ds, (float*)dh, &head);
  if (rc) return throw_wt(env, rc);
  napi_value out, hv;
  napi_create_object(env, &out);
  napi_create_int32(env, head, &hv);
  napi_set_named_property(env, out, "state", st);
  napi_set_named_property(env, out, "trail", tr);
  napi_set_named_property(env, out, "head", hv);
  return out;
}

// encodePost(h, derived) -> true when a statistics read was queued (SolverGPU.ts:225-261)
napi_value EncodePost(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  bool derived = true;
  napi_get_value_bool(env, argv[1], &derived);
  int32_t queued = 0;
  int rc = wt_encode_post(lk.h->s, derived ? 1 : 0, &queued);
  if (rc) return throw_wt(env, rc);
  napi_value out;
  napi_get_boolean(env, queued != 0, &out);
  return out;
}

// flushReads(h) -> {stats: [SolverStats of the completed reads, in posting order], pending}; never blocks
napi_value FlushReads(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  wt_stats st[8];
  int32_t count = 0, pending = 0;
  int rc = wt_flush_reads(lk.h->s, st, 8, &count, &pending);
  if (rc) return throw_wt(env, rc);
  napi_value out, arr, pv;
  napi_create_object(env, &out);
  napi_create_array_with_length(env, (size_t)count, &arr);
  for (int32_t k = 0; k < count; k++) napi_set_element(env, arr, (uint32_t)k, stats_object(env, st[k]));
  napi_create_int32(env, pending, &pv);
  napi_set_named_property(env, out, "stats", arr);
  napi_set_named_property(env, out, "pending", pv);
  return out;
}

// fence(h) -> ticket (a number): the point after all work queued on the handle so far
napi_value Fence(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  int64_t t = 0;
  int rc = wt_fence(lk.h->s, &t);
  if (rc) return throw_wt(env, rc);
  napi_value out;
  napi_create_double(env, (double)t, &out);
  return out;
}

// fenceDone(h, ticket) -> whether the GPU has passed the fence; never blocks
napi_value FenceDone(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!args(env, info, 2, argv)) return nullptr;
  HandleLock lk(env, argv[0]);
  if (!lk) return nullptr;
  double t = 0;
  napi_get_value_double(env, argv[1], &t);
  int32_t done = 0;
  int rc = wt_fence_done(lk.h->s, (int64_t)t, &done);
  if (rc) return throw_wt(env, rc);
  napi_value out;
  napi_get_boolean(env, done != 0, &out);
  return out;
}

napi_value BytesFor(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!args(env, info, 1, argv)) return nullptr;
  double nx = 0, ny = 0, nz = 0;
  get_prop_num(env, argv[0], "nx", &nx);
  get_prop_num(env, argv[0], "ny", &ny);
  get_prop_num(env, argv[0], "nz", &nz);
  wt_dims d{(int32_t)nx, (int32_t)ny, (int32_t)nz};
  napi_value out;
  napi_create_double(env, (double)wt_bytes_for(&d), &out);
  return out;
}

napi_value Init(napi_env env, napi_value exports) {
  struct { const char* name; napi_callback cb; } fns[] = {
      {"create", Create}, {"destroy", Destroy}, {"setParams", SetParams}, {"reset", Reset},
      {"clearStats", ClearStats}, {"sync", Sync}, {"info", Info}, {"uploadFlags", UploadFlags},
      {"readFlags", ReadFlags}, {"uploadSdf", UploadSdf}, {"setWheels", SetWheels}, {"step", Step},
      {"stepAndSample", StepAndSample}, {"stepAndSampleSync", StepAndSampleSync},
      {"readPopulations", ReadPopulations}, {"uploadPopulations", UploadPopulations},
      {"voxelizeSdf", VoxelizeSdf}, {"voxelizeMesh", VoxelizeMesh}, {"bytesFor", BytesFor},
      {"derive", Derive}, {"exportFields", ExportFields}, {"probe", Probe}, {"streamlines", Streamlines},
      {"particlesCreate", ParticlesCreate}, {"particlesDestroy", ParticlesDestroy},
      {"particlesUpdate", ParticlesUpdate}, {"particlesRead", ParticlesRead},
      {"encodePost", EncodePost}, {"flushReads", FlushReads}, {"fence", Fence}, {"fenceDone", FenceDone},
  };
  for (auto& f : fns) {
    napi_value fn;
    NAPI_CALL(env, napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn));
    NAPI_CALL(env, napi_set_named_property(env, exports, f.name, fn));
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
