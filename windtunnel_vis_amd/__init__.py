"""MI355X-native lattice-Boltzmann wind-tunnel solver (HIP kernels for gfx950 behind a C ABI).

The hot path — fused pull-stream + collide (D2Q9 / D3Q19, BGK / regularized, Smagorinsky LES),
interpolated walls, moving belt and rotating wheels, the voxelizers and the momentum-exchange
force reduction — lives in libwt.so (windtunnel_vis_amd/csrc, C ABI in include/wt.h). This
package is its Python host: `SolverGPU` mirrors the reference's solver surface.
"""
from .solver import BatchGPU, FieldExport, SolverGPU, SolverStats, Tracers, D2Q9, D3Q19  # noqa: F401

__all__ = ["BatchGPU", "FieldExport", "SolverGPU", "SolverStats", "D2Q9", "D3Q19"]
