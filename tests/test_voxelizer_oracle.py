"""Pins of the mesh-voxelizer restatement (oracle/vox_oracle.cpp) against closed forms and the
reference's own voxcheck bands (src/test/harness.ts:20-39)."""
import math

import numpy as np
import pytest

from oracle.oracle import decode_sdf, voxelize_mesh
from windtunnel_vis_amd import geometry as G

IDENT = np.eye(4, dtype=np.float32).T.reshape(-1)


def _box_count(lo, hi, n):
    k = np.arange(n) + 0.5
    return int(((k > lo) & (k < hi)).sum())


def test_box_exact_volume_and_frontal():
    dims = (48, 24, 20)
    lo, hi = (10.2, 5.3, 4.1), (30.7, 15.6, 12.9)
    m = G.box_mesh(lo, hi)
    flags, sdf, info = voxelize_mesh(m.positions, m.indices, IDENT, dims)
    cx, cy, cz = (_box_count(lo[a], hi[a], dims[a]) for a in range(3))
    assert info["solidCells"] == cx * cy * cz
    assert info["frontal"] == cy * cz
    f = flags.reshape(dims[2], dims[1], dims[0])
    z, y, x = np.meshgrid(*(np.arange(d) + 0.5 for d in dims[::-1]), indexing="ij")
    inside = (x > lo[0]) & (x < hi[0]) & (y > lo[1]) & (y < hi[1]) & (z > lo[2]) & (z < hi[2])
    assert np.array_equal(f != 0, inside)


def test_box_distances_are_exact_surface_distances():
    dims = (48, 24, 20)
    lo, hi = np.array([10.2, 5.3, 4.1]), np.array([30.7, 15.6, 12.9])
    m = G.box_mesh(lo, hi)
    flags, bits, info = voxelize_mesh(m.positions, m.indices, IDENT, dims)
    d = decode_sdf(bits).reshape(dims[2], dims[1], dims[0])
    z, y, x = np.meshgrid(*(np.arange(n) + 0.5 for n in dims[::-1]), indexing="ij")
    p = np.stack([x, y, z], -1)
    q = np.maximum(lo - p, p - hi)
    outside = np.linalg.norm(np.maximum(q, 0), axis=-1)
    inside = -q.max(axis=-1)
    exact = np.where(q.max(axis=-1) > 0, outside, inside)
    near = exact <= 2.5 - 1e-3
    far = exact > 2.5 + 1e-3
    assert np.all(d[near] >= 0)
    assert np.max(np.abs(d[near] - exact[near])) < 2e-5
    assert np.all(d[far] < 0)


@pytest.mark.parametrize("dims", [(128, 64, 64), (512, 192, 1)])
def test_sphere_voxcheck_bands(dims):
    """harness.ts:20-39: volume within 3 %, frontal area within 5 % of the analytic sphere."""
    sph = G.normalize_mesh(G.uv_sphere())
    M, _ = G.placement_matrix(sph, dims, mode="center", diam_frac=0.3, x_frac=0.35)
    flags, sdf, info = voxelize_mesh(sph.positions, sph.indices, M, dims)
    r = 0.3 * dims[1] / 2
    vol = 4 / 3 * math.pi * r ** 3 if dims[2] > 1 else math.pi * r * r
    area = math.pi * r * r if dims[2] > 1 else 2 * r
    assert abs(info["solidCells"] / vol - 1) < 0.03
    assert abs(info["frontal"] / area - 1) < 0.05


def test_overlapping_parts_union():
    """Two closed parts that overlap voxelize to their union (winding 2 is still inside)."""
    dims = (40, 20, 16)
    a = G.box_mesh((5.2, 3.1, 2.3), (20.6, 12.4, 10.8))
    b = G.box_mesh((15.3, 6.2, 5.1), (30.4, 16.7, 13.6))
    m = G.merge_meshes([a, b])
    flags, sdf, info = voxelize_mesh(m.positions, m.indices, IDENT, dims)
    fa, _, _ = voxelize_mesh(a.positions, a.indices, IDENT, dims)
    fb, _, _ = voxelize_mesh(b.positions, b.indices, IDENT, dims)
    assert np.array_equal(flags != 0, (fa != 0) | (fb != 0))


def test_thin_plate_kept_by_shell():
    """A plate thinner than a cell owns no cell centre; the shell keeps it (wgslVoxel.ts:126-199)."""
    dims = (32, 16, 16)
    m = G.box_mesh((8.2, 7.55, 3.3), (24.7, 7.85, 12.6))
    flags, sdf, info = voxelize_mesh(m.positions, m.indices, IDENT, dims)
    f = flags.reshape(dims[2], dims[1], dims[0])
    assert info["solidCells"] > 0
    assert set(np.nonzero(f)[1]) == {7}


def test_wheel_tags():
    dims = (40, 24, 20)
    body = G.box_mesh((8.2, 6.3, 5.1), (30.6, 14.4, 14.8))
    wheel = G.cylinder_mesh((12.0, 5.0, 4.0), 4.0, 1.6)
    m = G.merge_meshes([body, wheel])
    wheels = [dict(c=[12.0, 5.0, 4.0], axis=[0, 0, 1], r=4.0, hw=1.6)]
    flags, sdf, info = voxelize_mesh(m.positions, m.indices, IDENT, dims, wheels)
    tagged = (flags >> 8) == 1
    assert tagged.sum() > 0 and np.all(flags[tagged] & 1)
    z, y, x = np.meshgrid(*(np.arange(n) + 0.5 for n in dims[::-1]), indexing="ij")
    rad = np.hypot(x - 12.0, y - 5.0).reshape(-1)
    ax = np.abs(z - 4.0).reshape(-1)
    assert np.all(rad[tagged] <= 4.0 + 0.75 + 1e-6) and np.all(ax[tagged] <= 1.6 + 0.6 + 1e-6)
    solid = flags != 0
    assert np.all(tagged == (solid & (rad <= 4.75) & (ax <= 2.2)))
