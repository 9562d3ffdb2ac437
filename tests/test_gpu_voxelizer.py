"""GPU voxelizers. Mesh path: bit-exact against the CPU restatement (flags, distances, VoxelInfo)
and the reference's voxcheck bands. Analytic-SDF path: against the FP64 signed distance of the
same primitive union."""
import math

import numpy as np
import pytest

from oracle.oracle import voxelize_mesh
from windtunnel_vis_amd import geometry as G

pytestmark = pytest.mark.gpu

IDENT = np.eye(4, dtype=np.float32).T.reshape(-1)


def _solver(dims):
    from windtunnel_vis_amd import SolverGPU
    return SolverGPU(dims, dict(u=0.05, nu=0.01, cs=0.1, ground="noslip", spongeNu=0.0, spongeStart=0.9))


def _sdf_bits(s):
    d = s.readSdf()
    b = (~d.view(np.uint32)).astype(np.uint32)
    return np.where(d < 0, 0, b).astype(np.uint32)


def _meshes():
    sph = G.normalize_mesh(G.uv_sphere())
    out = []
    for dims in [(128, 64, 64), (512, 192, 1)]:
        M, _ = G.placement_matrix(sph, dims, mode="center", diam_frac=0.3, x_frac=0.35)
        out.append(("sphere%d" % dims[2], sph, M, dims, []))
    body = G.merge_meshes([G.box_mesh((8.2, 6.3, 5.1), (30.6, 14.4, 14.8)), G.cylinder_mesh((12.0, 5.0, 4.0), 4.0, 1.6),
                           G.box_mesh((8.2, 7.55, 16.3), (24.7, 7.85, 18.6))])
    out.append(("parts", body, IDENT, (40, 24, 20), [dict(c=[12.0, 5.0, 4.0], axis=[0, 0, 1], r=4.0, hw=1.6)]))
    yawed = G.normalize_mesh(G.merge_meshes([G.box_mesh((0, 0, 0), (3.0, 1.0, 1.2)), G.cylinder_mesh((0.6, 0.2, 0.1), 0.25, 0.1)]))
    M, _ = G.placement_matrix(yawed, (96, 48, 48), mode="ground", yaw_deg=12.0, pitch_deg=2.0, ride_cells=2.0)
    out.append(("yawed", yawed, M, (96, 48, 48), []))
    return out


@pytest.mark.parametrize("case", _meshes(), ids=lambda c: c[0])
def test_mesh_voxelizer_bit_exact(case):
    name, mesh, M, dims, wheels = case
    s = _solver(dims)
    info = s.voxelizeMesh(mesh.positions, mesh.indices, M, wheels)
    flags_ref, bits_ref, info_ref = voxelize_mesh(mesh.positions, mesh.indices, M, dims, wheels)
    flags = s.readFlags()
    assert info == info_ref, (info, info_ref)
    assert np.array_equal(flags, flags_ref), int((flags != flags_ref).sum())
    bits = _sdf_bits(s)
    assert np.array_equal(bits, bits_ref), int((bits != bits_ref).sum())
    s.destroy()


@pytest.mark.parametrize("dims", [(128, 64, 64), (512, 192, 1)])
def test_mesh_voxcheck_bands(dims):
    sph = G.normalize_mesh(G.uv_sphere())
    M, _ = G.placement_matrix(sph, dims, mode="center", diam_frac=0.3, x_frac=0.35)
    s = _solver(dims)
    info = s.voxelizeMesh(sph.positions, sph.indices, M)
    r = 0.3 * dims[1] / 2
    vol = 4 / 3 * math.pi * r ** 3 if dims[2] > 1 else math.pi * r * r
    area = math.pi * r * r if dims[2] > 1 else 2 * r
    assert abs(info["solidCells"] / vol - 1) < 0.03 and abs(info["frontal"] / area - 1) < 0.05


@pytest.mark.parametrize("maker", ["sedan", "truck", "wedge", "suv_yaw"])
def test_sdf_voxelizer_matches_fp64(maker):
    dims = {"sedan": (256, 128, 128), "truck": (512, 128, 128), "wedge": (320, 96, 128), "suv_yaw": (160, 80, 80)}[maker]
    body = {"sedan": lambda: G.sedan(dims), "truck": lambda: G.truck_trailer(dims),
            "wedge": lambda: G.sports_wedge(dims, length=200, width=64, height=30, corner=8),
            "suv_yaw": lambda: G.suv(dims, length=60, width=24, height=20, corner=5, wheel_r=5, wheel_w=4, yaw_deg=11.0)}[maker]()
    s = _solver(dims)
    s.setWheels(body.wheels)
    info = s.voxelizeSdf(body.prims())
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    d = body.sdf(np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1))
    flags = s.readFlags()
    solid = flags != 0
    # FP32 vs FP64: a sign may differ only within round-off of the surface
    mism = solid != (d < 0)
    assert np.all(np.abs(d[mism]) < 1e-4)
    assert info["solidCells"] == int(solid.sum())
    dist = s.readSdf()
    near = np.abs(d) < 2.5 - 1e-4
    assert np.all(dist[near] >= 0) and np.max(np.abs(dist[near] - np.abs(d[near]))) < 1e-4
    assert np.all(dist[np.abs(d) > 2.5 + 1e-4] < 0)
    if body.wheels:
        assert ((flags >> 8) > 0).sum() > 0
    s.destroy()


@pytest.mark.parametrize("pid", ["sedan", "sports", "suv", "truck", "f1", "ahmed25", "ahmed35", "sphere", "cylinder"])
def test_preset_meshes_bit_exact(pid):
    """The procedural presets through the mesh voxelizer, placed as the app places them (ground
    vehicles at a third of the tunnel length, validation bodies centred), wheels tagged."""
    from windtunnel_vis_amd import presets as P
    mesh = P.build_preset(pid)
    info_p = {p["id"]: p for p in P.PRESETS}[pid]
    dims = (128, 64, 64)
    mode = info_p["placement"]
    M, _ = G.placement_matrix(mesh, dims, mode=mode, ride_cells=1.0 if mode == "ground" else 0.0)
    wheels = G.grid_wheels(mesh, M)
    s = _solver(dims)
    s.setWheels(wheels)
    info = s.voxelizeMesh(mesh.positions, mesh.indices, M, wheels)
    flags_ref, bits_ref, info_ref = voxelize_mesh(mesh.positions, mesh.indices, M, dims, wheels)
    flags = s.readFlags()
    print(pid, info, "wheel cells", int(((flags >> 8) != 0).sum()), flush=True)
    assert info == info_ref
    assert np.array_equal(flags, flags_ref), int((flags != flags_ref).sum())
    assert np.array_equal(_sdf_bits(s), bits_ref)
    assert info["solidCells"] > 100 and info["frontal"] > 20
    if wheels:
        assert ((flags >> 8) != 0).any()
    st = s.stepAndSample(200)
    assert not st.unstable and math.isfinite(st.fx)
    s.destroy()
