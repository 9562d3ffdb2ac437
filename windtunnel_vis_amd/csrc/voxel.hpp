// Host-side interface of the voxelizer kernels.
#pragma once
#include "lbm.hpp"
#include "../../include/wt.h"

namespace wt {

constexpr int WT_MAX_PRIMS = 64;

struct VoxCtx {
  hipStream_t stream;
  int nx, ny, nz;
  int64_t n;
  uint32_t* flags;
  uint32_t* sdf;
  const KWheels* W;
};

// Writes flags and sdf for the whole grid; info8 = frontal, solid, min xyz, max xyz (inclusive).
int voxelize_sdf(const VoxCtx& v, const wt_prim* prims, int count, int64_t* info8);
int voxelize_mesh(const VoxCtx& v, const float* positions, int nv, const uint32_t* indices, int nt, const float* M16,
                  int64_t* info8);

}  // namespace wt
