// Host-side launch interface between the C-ABI layer and the kernels.
#pragma once
#include "lbm.hpp"

namespace wt {

struct LaunchCtx {
  hipStream_t stream;
  KParams P;
  int prec;                 // 0 f32, 1 f16, 2 bf16
  const uint32_t* flags;
  const uint32_t* sdf;
  uint8_t* kind;
  const KWheels* W;
};

void launch_init(const LaunchCtx& c, void* f);
void launch_classify(const LaunchCtx& c, uint32_t* list, uint32_t* masks, unsigned long long* counts);
// Links of the wall cells (direction-major, stride cap) and their meta words.
void launch_wallprep(const LaunchCtx& c, const uint32_t* list, const uint32_t* masks, WallLink* links, uint32_t* meta, int64_t count,
                     int64_t cap);
// chunk: items per XCD run (<= 0: one contiguous eighth of the items per XCD)
StepArgs step_args(const LaunchCtx& c, const uint32_t* list, const uint32_t* meta, const WallLink* links, int64_t count, int64_t cap,
                   int64_t chunk);
void launch_step(const LaunchCtx& c, const void* fin, void* fout, const StepArgs& A);
// One step of a batch of domains sharing lattice, storage, cells per thread and collision.
void launch_step_batch(hipStream_t st, int q, int prec, int cpt, bool reg, const DomainDesc* descs, const uint32_t* dmap,
                       uint32_t wall_total, uint32_t blocks, int parity);
int cells_per_thread(const KParams& P);
int64_t macro_blocks(int64_t n);
void launch_macro(const LaunchCtx& c, const void* fin, void* vel, void* meanA, float* meanB, double* partials, double* result);
void launch_unpack(const LaunchCtx& c, const void* f, float* out);
void launch_pack(const LaunchCtx& c, const float* in, void* f);
void launch_copy(hipStream_t st, const void* a, void* b, int64_t bytes);
void launch_clear_stats(const LaunchCtx& c, void* meanA, float* meanB);

}  // namespace wt
