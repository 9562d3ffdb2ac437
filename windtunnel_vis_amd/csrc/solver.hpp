// Private to libwt: the solver handle's state and the helpers the entry-point files share.
#pragma once
#include <hip/hip_runtime.h>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/wt.h"
#include "launch.hpp"
#include "mlp.hpp"

namespace wt {
// records the message for wt_last_error and returns the code
int fail(int code, const std::string& msg);
}  // namespace wt

#define HIP_TRY(call)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) return wt::fail(WT_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct wt_solver {
  // (the opaque handle of include/wt.h)
  int device = 0;
  hipStream_t stream = nullptr;
  wt_dims dims{};
  int64_t n = 0;
  int q = 0;
  int prec = 0;
  wt_params prm{};
  wt::KParams P{};
  wt::KWheels W{};
  void* f[2] = {nullptr, nullptr};
  uint32_t* flags = nullptr;
  uint32_t* sdf = nullptr;
  uint8_t* kind = nullptr;
  uint32_t* wlist = nullptr;   // cells handled by the wall kernel
  uint32_t* wmask = nullptr;   // their solid masks
  wt::WallLink* wlink = nullptr;   // their links, (q - 1) x wcap
  uint32_t* wmeta = nullptr;   // their meta words
  int64_t wcap = 0;
  unsigned long long* counts = nullptr;
  ushort4* vel = nullptr;
  float4* meanA = nullptr;
  float* meanB = nullptr;
  double* partials = nullptr;
  // derived fields (allocated on first use): (wx, wy, wz, Q), (TI, solid, 0, 0), mean (ux, uy, uz, rho)
  ushort4* vort = nullptr;
  ushort4* stat = nullptr;
  ushort4* meanh = nullptr;
  double* result = nullptr;
  double* h_result = nullptr;  // pinned
  int64_t nwall = 0, nfast = 0, nsolid = 0;  // cells: wall kernel, bulk fluid, bulk solid
  int parity = 0;
  bool pending_init = true;
  int64_t step = 0;
  int64_t bytes = 0;
  // profiling
  int prof = 0;                 // WT_PROFILE_* mode
  std::vector<hipEvent_t> ev;   // event pairs: per launch (WT_PROFILE_LAUNCH) or per step call (WT_PROFILE_SPAN)
  std::vector<int64_t> ev_launches;  // launches covered by each pair
  int item_chunk = 4;           // WT_TUNE_ITEM_CHUNK (4: plain order, the fastest untuned choice; 0: eighths)
  int cpt = 2;                  // WT_TUNE_CELLS_PER_THREAD (2 needs an even nx)
  size_t ev_used = 0;
  double acc_bulk = 0, acc_bnd = 0;
  int64_t acc_launches = 0;
  // learned closure
  uint8_t* wdist = nullptr;     // wall distance per cell (cells x 32, capped at 4 cells); built on demand
  bool wdist_valid = false;
  wt::MlpDev* mlp = nullptr;        // network in the MFMA layout (device)
  int64_t* pull = nullptr;          // bulk pull table (device), rebuilt with the parameters
  std::vector<int64_t> h_pull;      // its host copy
  // placement of the wall blocks among the bulk blocks (StepArgs::role): the row of the first cell
  // of every group of 8 x BLOCK wall cells (the wall list is sorted by cell), and the device table
  // with the launch layout it was built for
  std::vector<uint32_t> h_wrows;
  uint64_t wall_version = 0;        // bumped whenever the wall list is rebuilt
  uint32_t* role = nullptr;
  size_t role_cap = 0;
  uint64_t role_key[6] = {};
  // asynchronous statistics reads (wt_encode_post / wt_flush_reads): a ring of pinned slots, each
  // filled by a copy on the handle's stream and marked by an event, like the reference's staging
  // buffers (SolverGPU.ts:86-88, :249-257)
  struct ReadSlot {
    double* host = nullptr;  // pinned, 8 doubles
    hipEvent_t ev = nullptr;
    bool busy = false;
    int64_t step = 0;
    uint64_t seq = 0;        // order of the posts
    uint32_t gen = 0;        // bumped at every reuse (asynchronous samples' tickets)
  };
  std::vector<ReadSlot> reads;
  uint64_t read_seq = 0;
  // asynchronous stepAndSample (wt_step_and_sample_async / wt_sample_wait): a pool that grows
  // when every slot is in flight; `read_mu` guards it, since wt_sample_wait may run on another thread
  std::vector<ReadSlot*> samples;
  std::mutex read_mu;
  // queue fences (wt_fence): a ring of events; fence_ticket[k] is the ticket last recorded in slot
  // k, fence_done the newest ticket known to be passed (in stream order every older one is too)
  static constexpr int FENCES = 16;
  hipEvent_t fence_ev[FENCES] = {};
  int64_t fence_ticket[FENCES] = {};
  int64_t fence_seq = 0, fence_done = 0;

  wt::LaunchCtx ctx() const {
    wt::LaunchCtx c;
    c.stream = stream;
    c.P = P;
    c.prec = prec;
    c.flags = flags;
    c.sdf = sdf;
    c.kind = kind;
    c.W = &W;
    c.wdist = wdist;
    c.mlp = mlp;
    c.pull = pull;
    return c;
  }
};

namespace wt {
void derive(wt_solver* s);
int run_init(wt_solver* s);
int prep_walls(wt_solver* s);
int reclassify(wt_solver* s);
int ensure_roles(wt_solver* s, wt::StepArgs& A);
int do_steps(wt_solver* s, int steps);
int do_sample(wt_solver* s);
void decode_stats(const wt_solver* s, wt_stats* out);
void decode_result(const double* d, int64_t step, wt_stats* out);
int64_t storage_bytes(const wt_solver* s);  // bytes of one population buffer
int ensure_wdist(wt_solver* s);             // wall distances for the learned closure
}  // namespace wt
