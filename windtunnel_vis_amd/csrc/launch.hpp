// Host-side launch interface between the C-ABI layer and the kernels.
#pragma once
#include "../../include/wt.h"
#include "lbm.hpp"
#include "mlp.hpp"
#include <vector>

namespace wt {

struct LaunchCtx {
  hipStream_t stream;
  KParams P;
  int prec;                 // 0 f32, 1 f16, 2 bf16
  const uint32_t* flags;
  const uint32_t* sdf;
  uint8_t* kind;
  const KWheels* W;
  const uint8_t* wdist;     // wall distance bytes (learned closure), may be null
  const MlpDev* mlp;        // learned closure network (device), null: Smagorinsky
  const int64_t* pull;      // pull table (device), see pull_source
};

// The bulk pull table of a parameter set: pull_table_len(q) offsets (host memory).
int pull_table_len(int q);
void pull_table(const KParams& P, int64_t* out);

void launch_init(const LaunchCtx& c, void* f);
void launch_classify(const LaunchCtx& c, uint32_t* list, uint32_t* masks, unsigned long long* counts);
// Links of the wall cells (direction-major, stride cap) and their meta words.
void launch_wallprep(const LaunchCtx& c, const uint32_t* list, const uint32_t* masks, WallLink* links, uint32_t* meta, int64_t count,
                     int64_t cap);
// chunk: items per XCD run (<= 0: one contiguous eighth of the items per XCD)
// Role table of a step launch (StepArgs::role) for the layout in A: wall group k (8 x BLOCK wall
// cells, the first in row wall_rows[k]) goes right after the bulk group that updates row
// wall_rows[k] + ny + 1, by when the bulk blocks have streamed the rows its links read (the cell's
// own row and its y / z neighbours).
constexpr int WALL_GROUP_CELLS = 8 * 256;  // wall cells per group of 8 wall blocks (the step kernels' 256-thread blocks)
std::vector<uint32_t> build_roles(const StepArgs& A, const std::vector<uint32_t>& wall_rows, int ny, int nz);
StepArgs step_args(const LaunchCtx& c, const uint32_t* list, const uint32_t* meta, const WallLink* links, int64_t count, int64_t cap,
                   int64_t chunk);
// ev0 / ev1 (may be null): events recorded at the step kernel's own start and end
void launch_step(const LaunchCtx& c, const void* fin, void* fout, const StepArgs& A, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// One step of a batch of domains sharing lattice, storage, cells per thread and collision.
void launch_step_batch(hipStream_t st, int q, int prec, int cpt, bool reg, const DomainDesc* descs, const uint32_t* dmap,
                       uint32_t blocks, int parity);
int cells_per_thread(const KParams& P);
int64_t macro_records(int64_t n);  // 8-double records of the macro pass's partials buffer
void launch_macro(const LaunchCtx& c, const void* fin, void* vel, void* meanA, float* meanB, double* partials, double* result);
void launch_unpack(const LaunchCtx& c, const void* f, float* out);
void launch_unpack_planes(const LaunchCtx& c, const void* f, int32_t z0, int32_t count, float* out);
void launch_pack(const LaunchCtx& c, const float* in, void* f);
void launch_copy(hipStream_t st, const void* a, void* b, int64_t bytes);
void launch_clear_stats(const LaunchCtx& c, void* meanA, float* meanB);

// ---- derived fields and export (kernels_fields.hip); fields are 4 halves per cell ----
void launch_derived(const LaunchCtx& c, const void* vel, const void* meanA, const float* meanB, void* vort, void* stat, void* mean);
void launch_export(hipStream_t st, const void* vel, const void* mean, const void* vort, int nx, int ny, int nz, int stride,
                   const int (&o)[3], float* ov, float* om, float* ow);
void launch_unpack_h4(hipStream_t st, const void* in, int64_t n, float* out);
// probes (2 float4 per point: field, mean field), streamlines (points float4 per seed), tracers
void launch_probes(hipStream_t st, const void* vel, const void* mean, const uint8_t* kind, int nx, int ny, int nz, const float* pos,
                   int count, float* out);
void launch_streamlines(hipStream_t st, const void* field, const uint8_t* kind, int nx, int ny, int nz, const wt_rake& rake, int seeds,
                        int per_line, int points, float h, float U, float* out);
void launch_particles(hipStream_t st, const void* vel, const uint8_t* kind, int nx, int ny, int nz, const wt_particle_opts& o, int head,
                      float steps, uint32_t seed, int fill, int advance, float* state, float* hist);

// ---- learned closure (kernels_mlp.hip) ----
void launch_wall_distance(hipStream_t st, const uint32_t* flags, uint32_t* tmp, uint8_t* out, int nx, int ny, int nz);
int64_t coarse_blocks(int64_t nc);
void launch_coarse_fields(hipStream_t st, int prec, const void* f, const uint32_t* flags, int fnx, int fny, int fnz, float* cf,
                          uint8_t* bad);
void launch_coarse_count(hipStream_t st, const float* cf, const uint8_t* bad, float2* nd, const uint8_t* wd, int cnx, int cny, int cnz,
                         float nu0, uint32_t* counts);
void launch_coarse_samples(hipStream_t st, const float* cf, const uint8_t* bad, const float2* nd, const uint8_t* wd, int cnx, int cny,
                           int cnz, float nu0, const uint64_t* offsets, float* out);
int64_t grad_partial_floats(int64_t samples);
void launch_mlp_grad(hipStream_t st, const float* data, const uint32_t* perm, int64_t first, int64_t count, float inv_batch,
                     const float* params, const MlpNorm& norm, float* partials, float* grad, double* lossacc);
void launch_adam(hipStream_t st, float* p, float* m, float* v, const float* g, float lr, float b1, float b2, float eps, int64_t t);
void launch_mlp_infer(hipStream_t st, const MlpDev* net, const float* raw, int64_t n, float* nu);

}  // namespace wt
