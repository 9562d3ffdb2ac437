"""CSV / VTK writers (src/analysis/export.ts) and the derived-field oracle (wgslLbm.ts:527-592),
on the CPU. Number printing is pinned against JavaScript's own Number formatting
(tests/golden/js_number_vectors.json, written by make_js_number_vectors.js with Node)."""
import json
import os
import struct

import numpy as np
import pytest

from oracle import fields_oracle as FO
from windtunnel_vis_amd import export as E
from windtunnel_vis_amd.solver import FieldExport

GOLD = os.path.join(os.path.dirname(__file__), "golden", "js_number_vectors.json")


def _vectors():
    with open(GOLD) as fh:
        return json.load(fh)


def _double(bits):
    return struct.unpack(">d", bytes.fromhex(bits))[0]


def test_js_number_matches_node():
    for v in _vectors():
        x = _double(v["bits"])
        assert E.js_number(x) == v["str"], (x, v["str"])


def test_js_to_precision_matches_node():
    for v in _vectors():
        x = _double(v["bits"])
        assert E.js_number(E.js_to_precision(x, 7)) == v["p7"], (x, v["p7"])


def test_float32_big_endian_matches_dataview():
    vals = [_double(v["bits"]) for v in _vectors()]
    with np.errstate(over="ignore"):
        got = np.asarray(vals, dtype=np.float64).astype(">f4").tobytes()
    want = b"".join(bytes.fromhex(v["f32be"]) for v in _vectors())
    assert got == want


def test_csv_layout():
    txt = E.csv(["t", "cd", "label"], [[1, 0.123456789, "a"], [2, float("nan"), "b"], [1234567.5, float("inf"), "c"]],
                comments=["run 1", "u=0.06"])
    assert txt == "# run 1\n# u=0.06\nt,cd,label\n1,0.1234568,a\n2,,b\n1234568,,c\n"


def test_stride_for():
    assert E.stride_for(1000, 1_000_000) == 1
    assert E.stride_for(8_000_000, 1_000_000) == 2
    assert E.stride_for(512 * 256 * 256, 1_000_000) == 4  # cbrt(33.55) = 3.22
    assert E.stride_for(8_000_001, 1_000_000) == 3


def _field(dims, stride, seed=1):
    rng = np.random.default_rng(seed)
    m = dims[0] * dims[1] * dims[2]
    vel = rng.normal(0.05, 0.02, (m, 4)).astype(np.float32)
    vel[:, 3] = 1 + rng.normal(0, 1e-3, m)
    mean = rng.normal(0.05, 0.01, (m, 4)).astype(np.float32)
    mean[:, 3] = 1 + rng.normal(0, 1e-3, m)
    vort = rng.normal(0, 1e-3, (m, 4)).astype(np.float32)
    return FieldExport(dims, stride, vel, mean, vort)


@pytest.mark.parametrize("dims,stride", [((5, 4, 3), 2), ((7, 6, 1), 1)])
def test_vtk_round_trip(dims, stride):
    f = _field(dims, stride)
    U, L, dx, rho_ref = 0.06, 48.0, 0.0125, 1.0003
    data = E.vtk(f, dx=dx, U=U, L=L, rhoRef=rho_ref, title="truck | 512x256x256")
    hdr, arr = E.read_vtk(data)
    assert data.startswith(b"# vtk DataFile Version 3.0\ntruck | 512x256x256\nBINARY\nDATASET STRUCTURED_POINTS\n")
    assert hdr["dims"] == tuple(dims)
    assert hdr["origin"] == [E.js_number(0.5 * dx)] * 2 + [E.js_number(0.5 * dx) if dims[2] > 1 else "0"]
    assert hdr["spacing"] == [E.js_number(dx * stride)] * 3
    # values: computed in double from the float32 inputs, stored as float32 (DataView.setFloat32)
    v64 = f.vel.astype(np.float64)
    m64 = f.mean.astype(np.float64)
    w64 = f.vort.astype(np.float64)
    np.testing.assert_array_equal(arr["velocity"], (v64[:, :3] / U).astype(np.float32))
    np.testing.assert_array_equal(arr["mean_velocity"], (m64[:, :3] / U).astype(np.float32))
    np.testing.assert_array_equal(arr["vorticity"], (w64[:, :3] * L / U).astype(np.float32))
    np.testing.assert_array_equal(arr["Cp"], (2 * (v64[:, 3] - rho_ref) / (3 * U * U)).astype(np.float32))
    np.testing.assert_array_equal(arr["mean_Cp"], (2 * (m64[:, 3] - rho_ref) / (3 * U * U)).astype(np.float32))
    np.testing.assert_array_equal(arr["Q_criterion"], (w64[:, 3] * L * L / (U * U)).astype(np.float32))
    assert list(arr) == ["velocity", "mean_velocity", "vorticity", "Cp", "mean_Cp", "Q_criterion"]
    # big-endian: the first velocity component read back by hand
    off = data.index(b"VECTORS velocity float\n") + len(b"VECTORS velocity float\n")
    assert struct.unpack(">f", data[off:off + 4])[0] == np.float32(v64[0, 0] / U)


def test_vtk_title_truncated():
    f = _field((2, 2, 2), 1)
    data = E.vtk(f, dx=1.0, U=0.1, L=10.0, rhoRef=1.0, title="x" * 300)
    assert data.split(b"\n")[1] == b"x" * 250


# ---- derived-field oracle pins (closed forms) ----------------------------------------------------

def _grid(dims):
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return x.reshape(-1).astype(float), y.reshape(-1).astype(float), z.reshape(-1).astype(float)


def _interior(dims):
    nx, ny, nz = dims
    x, y, z = _grid(dims)
    ok = (x > 0) & (x < nx - 1) & (y > 0) & (y < ny - 1)
    if nz > 1:
        ok &= (z > 0) & (z < nz - 1)
    return ok


def _run(u, dims, solid=None):
    n = len(u)
    vel = np.concatenate([u, np.ones((n, 1))], axis=1)
    solid = np.zeros(n, bool) if solid is None else solid
    return FO.derived(vel, solid, dims, np.zeros((n, 4)), np.zeros(n), 0.05)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_rigid_rotation(axis):
    dims = (6, 5, 7)
    x, y, z = _grid(dims)
    om = 0.01
    r = [x - 3, y - 2, z - 3]
    a, b = [(1, 2), (2, 0), (0, 1)][axis]  # rotation in the (a, b) plane about `axis`
    u = np.zeros((len(x), 3))
    u[:, a] = -om * r[b]
    u[:, b] = om * r[a]
    vort, _, _ = _run(u, dims)
    ok = _interior(dims)
    w = np.zeros(3)
    w[axis] = 2 * om
    np.testing.assert_allclose(vort[ok, :3], np.broadcast_to(w, (ok.sum(), 3)), atol=1e-15)
    np.testing.assert_allclose(vort[ok, 3], om * om, rtol=1e-12)  # Q = |Omega|^2 / 2 = om^2


def test_pure_shear_and_strain():
    dims = (6, 5, 4)
    x, y, z = _grid(dims)
    g = 0.02
    u = np.stack([g * y, 0 * x, 0 * x], axis=1)  # shear: wz = -g, Q = 0
    vort, _, _ = _run(u, dims)
    ok = _interior(dims)
    np.testing.assert_allclose(vort[ok, 2], -g, rtol=1e-12)
    np.testing.assert_allclose(vort[ok, 3], 0.0, atol=1e-18)
    u = np.stack([g * x, -g * y, 0 * x], axis=1)  # plane strain: w = 0, Q = -g^2
    vort, _, _ = _run(u, dims)
    np.testing.assert_allclose(vort[ok, :3], 0.0, atol=1e-18)
    np.testing.assert_allclose(vort[ok, 3], -g * g, rtol=1e-12)


def test_clamped_face_is_half_one_sided():
    dims = (5, 4, 3)
    x, y, z = _grid(dims)
    u = np.stack([0 * x, 0.01 * x, 0 * x], axis=1)  # wz = d uy / dx = 0.01 inside
    vort, _, _ = _run(u, dims)
    face = (x == 0) | (x == dims[0] - 1)
    np.testing.assert_allclose(vort[face, 2], 0.005, rtol=1e-12)
    np.testing.assert_allclose(vort[~face, 2], 0.01, rtol=1e-12)


def test_near_wall_fade():
    dims = (9, 9, 9)
    x, y, z = _grid(dims)
    om = 0.01
    u = np.stack([-om * (y - 4), om * (x - 4), 0 * x], axis=1)
    solid = (x == 4) & (y == 4) & (z == 4)
    vort, stat, _ = _run(u, dims, solid)
    q = vort[:, 3].reshape(9, 9, 9)  # (z, y, x)
    assert q[4, 4, 4] == 0 and stat[solid, 1][0] == 1.0
    assert q[4, 4, 5] == 0 and q[4, 3, 4] == 0 and q[5, 4, 4] == 0  # axis distance 1: off
    np.testing.assert_allclose([q[4, 4, 6], q[4, 2, 4], q[2, 4, 4]], 0.5 * om * om, rtol=1e-12)  # distance 2: half
    np.testing.assert_allclose([q[4, 5, 5], q[5, 5, 4], q[4, 4, 7]], om * om, rtol=1e-12)  # diagonal / far: full


def test_turbulence_intensity_identity():
    n = 64
    rng = np.random.default_rng(3)
    m = rng.normal(0, 0.05, (n, 4))
    kk = rng.uniform(0, 1e-4, n)
    mb = (m[:, :3] ** 2).sum(1) + kk
    mb[:5] = (m[:5, :3] ** 2).sum(1) - 1e-6  # negative variance clamps to 0
    solid = np.zeros(n, bool)
    solid[10] = True
    _, stat, mean = FO.derived(np.zeros((n, 4)), solid, (n, 1, 1), m, mb, 0.06)
    want = np.sqrt(kk / 3) / 0.06
    want[:5] = 0
    want[10] = 0
    np.testing.assert_allclose(stat[:, 0], want, rtol=1e-9, atol=1e-12)
    np.testing.assert_array_equal(stat[:, 1], solid.astype(float))
    np.testing.assert_array_equal(mean, m)


@pytest.mark.parametrize("dims", [(5, 4, 3), (6, 5, 1)])
def test_vectorised_matches_loops(dims):
    rng = np.random.default_rng(7)
    n = dims[0] * dims[1] * dims[2]
    vel = rng.normal(0, 0.05, (n, 4))
    solid = rng.random(n) < 0.15
    ma = rng.normal(0, 0.05, (n, 4))
    mb = rng.uniform(0, 0.01, n)
    v1, s1, _ = FO.derived(vel, solid, dims, ma, mb, 0.05)
    v2, s2 = FO.derived_loops(vel, solid, dims, ma, mb, 0.05)
    np.testing.assert_allclose(v1, v2, rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(s1, s2, rtol=1e-12, atol=1e-18)


def test_export_sampling():
    dims = (7, 5, 4)
    n = 7 * 5 * 4
    f = np.arange(4 * n, dtype=float).reshape(n, 4)
    np.testing.assert_array_equal(FO.export_sample(f, dims, 1), f)
    s = FO.export_sample(f, dims, 3)
    assert FO.export_dims(dims, 3) == (3, 2, 2)
    want = []
    for gz in range(2):
        for gy in range(2):
            for gx in range(3):
                px, py, pz = min(gx * 3, 6), min(gy * 3, 4), min(gz * 3, 3)
                want.append(f[px + 7 * (py + 5 * pz)])
    np.testing.assert_array_equal(s, np.array(want))
