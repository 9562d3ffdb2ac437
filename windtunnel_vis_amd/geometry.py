"""Host-side geometry: analytic obstacle flags (the reference's roundFlags) and the procedural
analytic-SDF bodies the workloads are defined on (sedan / SUV, truck + trailer, sports wedge),
expressed as primitive lists for the GPU voxelizer (wt_voxelize_sdf).

Placement follows the reference's ground mode (src/voxelize/mesh.ts:354-379): x centre at
xFrac * nx, z centre at nz / 2, lowest point at y = ride height, yaw about y and pitch about z
through the body centre.
"""
import math

import numpy as np

from . import _native as N


def round_flags(dims, c, r, cylinder):
    """Analytic circle (2D) / sphere (3D) flags and wall distances (validation.ts:35-53).

    Returns (flags u32[n], frontal, sdf f32[n] with -1 = unknown)."""
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    dx = x + 0.5 - c[0]
    dy = y + 0.5 - c[1]
    dz = np.zeros_like(dx, dtype=np.float64) if cylinder else z + 0.5 - c[2]
    r2 = dx * dx + dy * dy + dz * dz
    dist = np.abs(np.sqrt(r2) - r)
    sdf = np.where(dist < 2.5, dist, -1.0).astype(np.float32).reshape(-1)
    solid = r2 < r * r
    flags = solid.astype(np.uint32).reshape(-1)
    proj = solid.any(axis=2)  # over x -> (z, y)
    return flags, int(proj.sum()), sdf


def rot_y(deg):
    a = math.radians(deg)
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def rot_z(deg):
    a = math.radians(deg)
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def frame_for_axis(axis):
    """Rotation whose third column is `axis` (the cylinder's local z)."""
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    t = np.array([1.0, 0, 0]) if abs(a[0]) < 0.9 else np.array([0, 1.0, 0])
    e1 = np.cross(t, a)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(a, e1)
    return np.stack([e1, e2, a], axis=1)


def prim(kind, center, rot=None, half=(0, 0, 0), radius=0.0, param=0.0):
    p = N.Prim()
    p.kind = kind
    p.flags = 0
    R = np.eye(3) if rot is None else np.asarray(rot, dtype=np.float64)
    for a in range(3):
        p.center[a] = float(center[a])
        p.half[a] = float(half[a])
    for k in range(9):
        p.rot[k] = float(R.reshape(-1)[k])
    p.radius = float(radius)
    p.param = float(param)
    return p


class Body:
    """A union of analytic primitives plus its rotating wheels, in grid space."""

    def __init__(self):
        self.parts = []  # (kind, center, R, half, radius, param)
        self.wheels = []  # dicts {c, axis, r, hw}
        self.length = 0.0

    def add(self, kind, center, R=None, half=(0, 0, 0), radius=0.0, param=0.0):
        self.parts.append([kind, np.asarray(center, dtype=np.float64), np.eye(3) if R is None else np.asarray(R, dtype=np.float64),
                           np.asarray(half, dtype=np.float64), float(radius), float(param)])

    def add_wheel(self, center, axis, r, hw):
        self.add(N.PRIM_CYLINDER, center, frame_for_axis(axis), (r, hw, 0))
        self.wheels.append(dict(c=list(map(float, center)), axis=list(map(float, axis)), r=float(r), hw=float(hw)))

    def transform(self, R, pivot):
        """Rotate every part and wheel by R about `pivot`."""
        pivot = np.asarray(pivot, dtype=np.float64)
        for p in self.parts:
            p[1] = pivot + R @ (p[1] - pivot)
            p[2] = R @ p[2]
        for w in self.wheels:
            w["c"] = list(map(float, pivot + R @ (np.asarray(w["c"]) - pivot)))
            w["axis"] = list(map(float, R @ np.asarray(w["axis"])))

    def translate(self, d):
        d = np.asarray(d, dtype=np.float64)
        for p in self.parts:
            p[1] = p[1] + d
        for w in self.wheels:
            w["c"] = list(map(float, np.asarray(w["c"]) + d))

    def prims(self):
        return [prim(k, c, R, h, r, q) for k, c, R, h, r, q in self.parts]

    def sdf(self, pts):
        """Signed distance at points (m, 3) on the CPU (FP64) — the voxelizer's formula."""
        d = np.full(len(pts), np.inf)
        for kind, c, R, h, r, q in self.parts:
            l = (pts - c) @ R  # R^T (p - c) for each row
            if kind == N.PRIM_SPHERE:
                dk = np.linalg.norm(l, axis=1) - h[0]
            elif kind == N.PRIM_CYLINDER:
                a = np.hypot(l[:, 0], l[:, 1]) - h[0]
                b = np.abs(l[:, 2]) - h[1]
                dk = np.minimum(np.maximum(a, b), 0) + np.hypot(np.maximum(a, 0), np.maximum(b, 0))
            else:
                rr = min(r, h.min())
                qv = np.abs(l) - (h - rr)
                dk = np.linalg.norm(np.maximum(qv, 0), axis=1) + np.minimum(qv.max(axis=1), 0) - rr
                if kind == N.PRIM_WEDGE:
                    plane = (l[:, 1] - (h[1] - q * (h[0] - l[:, 0]))) / math.sqrt(1 + q * q)
                    dk = np.maximum(dk, plane)
            d = np.minimum(d, dk)
        return d

    def lowest_y(self):
        """Lowest point of the body (sampled on each part's surface bounds; exact for the parts used)."""
        lo = math.inf
        for kind, c, R, h, r, q in self.parts:
            if kind == N.PRIM_CYLINDER:
                ax = R[:, 2]
                # lowest point of a cylinder with axis a, radius rc, half width hw
                rc, hw = h[0], h[1]
                ext = rc * math.sqrt(max(0.0, 1 - ax[1] ** 2)) + hw * abs(ax[1])
                lo = min(lo, c[1] - ext)
            elif kind == N.PRIM_SPHERE:
                lo = min(lo, c[1] - h[0])
            else:
                ext = sum(abs(R[1, k]) * h[k] for k in range(3))
                lo = min(lo, c[1] - ext)
        return lo


def _place(body, dims, x_frac, ride, yaw_deg, pitch_deg):
    nx, ny, nz = dims
    pivot = np.array([0.0, 0.0, 0.0])
    body.transform(rot_y(yaw_deg) @ rot_z(-pitch_deg), pivot)
    zc = nz / 2 if nz > 1 else 0.5
    lo = body.lowest_y()
    body.translate([x_frac * nx, -lo + ride, zc])
    return body


def sedan(dims, length=85, width=32, height=22, corner=6, wheel_r=6, wheel_w=5, ride=3, x_frac=0.34, yaw_deg=0.0,
          pitch_deg=0.0):
    """Procedural sedan: a rounded box L x H x W (x, y, z) with corner radius `corner`, its bottom at the
    ride height, and four lateral-axis wheel cylinders touching the floor at 0.165 L and 0.79 L from the
    nose (the reference's sedan proportions, presets.ts:205-215)."""
    b = Body()
    b.length = float(length)
    # body centred at the origin in x/z; y measured from the wheel contact line (y = 0)
    b.add(N.PRIM_ROUNDED_BOX, (0, ride + height / 2, 0), None, (length / 2, height / 2, width / 2), corner)
    zoff = width / 2 - wheel_w / 2 - 0.5
    for fx in (0.165, 0.79):
        for s in (-1, 1):
            b.add_wheel((-length / 2 + fx * length, wheel_r, s * zoff), (0, 0, 1), wheel_r, wheel_w / 2)
    # lowest point is the wheel bottom at y = 0, so placement puts the wheels on the floor and the
    # body bottom at `ride`
    return _place(b, dims, x_frac, 0.0, yaw_deg, pitch_deg)


def suv(dims, length=85, width=34, height=30, corner=7, wheel_r=7, wheel_w=6, ride=4, x_frac=0.34, yaw_deg=0.0,
        pitch_deg=0.0):
    """Taller SUV variant of the sedan (SUV proportions, presets.ts:233-244)."""
    return sedan(dims, length, width, height, corner, wheel_r, wheel_w, ride, x_frac, yaw_deg, pitch_deg)


def truck_trailer(dims, cab=(60, 28, 24), trailer=(110, 30, 26), gap=6, corner=4, wheel_r=6, wheel_w=5, ride=3,
                  x_frac=0.3, yaw_deg=0.0):
    """Truck + trailer: two rounded boxes (L x H x W) separated by `gap`, ten wheels."""
    b = Body()
    Lc, Hc, Wc = cab
    Lt, Ht, Wt = trailer
    total = Lc + gap + Lt
    b.length = float(total)
    x0 = -total / 2
    b.add(N.PRIM_ROUNDED_BOX, (x0 + Lc / 2, ride + Hc / 2, 0), None, (Lc / 2, Hc / 2, Wc / 2), corner)
    b.add(N.PRIM_ROUNDED_BOX, (x0 + Lc + gap + Lt / 2, ride + Ht / 2, 0), None, (Lt / 2, Ht / 2, Wt / 2), corner)
    xs = [x0 + 0.2 * Lc, x0 + 0.75 * Lc, x0 + Lc + gap + 0.7 * Lt, x0 + Lc + gap + 0.8 * Lt, x0 + Lc + gap + 0.9 * Lt]
    for x in xs:
        W = Wc if x < x0 + Lc else Wt
        zoff = W / 2 - wheel_w / 2 - 0.5
        for s in (-1, 1):
            b.add_wheel((x, wheel_r, s * zoff), (0, 0, 1), wheel_r, wheel_w / 2)
    return _place(b, dims, x_frac, 0.0, yaw_deg, 0.0)


def sports_wedge(dims, length=300, width=90, height=45, corner=10, ride=3, x_frac=0.34, slope=0.12):
    """Low rounded wedge (sports car): a rounded box whose top slopes down toward the nose."""
    b = Body()
    b.length = float(length)
    b.add(N.PRIM_WEDGE, (0, ride + height / 2, 0), None, (length / 2, height / 2, width / 2), corner, slope)
    return _place(b, dims, x_frac, ride, 0.0, 0.0)


def sphere_body(dims, center, radius):
    b = Body()
    b.length = 2.0 * radius
    b.add(N.PRIM_SPHERE, center, None, (radius, 0, 0))
    return b


# ---------------------------------------------------------------------------------------------
# Triangle meshes for the mesh voxelizer (wt_voxelize_mesh).
# ---------------------------------------------------------------------------------------------
class Mesh:
    """positions (V, 3) float32, indices (T, 3) uint32, optional wheels (mesh space), bbox."""

    def __init__(self, positions, indices, wheels=None):
        self.positions = np.asarray(positions, dtype=np.float32).reshape(-1, 3)
        self.indices = np.asarray(indices, dtype=np.uint32).reshape(-1, 3)
        self.wheels = wheels or []

    @property
    def min(self):
        return self.positions.min(axis=0).astype(np.float64)

    @property
    def max(self):
        return self.positions.max(axis=0).astype(np.float64)


def uv_sphere(radius=0.5, width_segments=96, height_segments=64):
    """Latitude-longitude sphere with outward-oriented triangles (the sphere preset is a 96 x 64
    UV sphere of radius 0.5, presets.ts:366-368). Pole rows emit one triangle per quad."""
    verts = []
    for iy in range(height_segments + 1):
        v = iy / height_segments
        th = v * math.pi
        for ix in range(width_segments + 1):
            u = ix / width_segments
            ph = u * 2 * math.pi
            verts.append((-radius * math.cos(ph) * math.sin(th), radius * math.cos(th), radius * math.sin(ph) * math.sin(th)))
    W = width_segments + 1
    tris = []
    for iy in range(height_segments):
        for ix in range(width_segments):
            a = iy * W + ix + 1
            b = iy * W + ix
            c = (iy + 1) * W + ix
            d = (iy + 1) * W + ix + 1
            if iy != 0:
                tris.append((a, b, d))
            if iy != height_segments - 1:
                tris.append((b, c, d))
    return Mesh(verts, tris)


def box_mesh(lo, hi):
    """Axis-aligned box with outward-oriented triangles."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = [(x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0), (x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)]
    f = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4), (3, 7, 6), (3, 6, 2), (0, 4, 7), (0, 7, 3),
         (1, 2, 6), (1, 6, 5)]
    return Mesh(v, f)


def cylinder_mesh(center, radius, half_width, segments=48):
    """Closed cylinder with axis z (a wheel), outward-oriented."""
    cx, cy, cz = center
    v = []
    for k in range(segments):
        a = 2 * math.pi * k / segments
        v.append((cx + radius * math.cos(a), cy + radius * math.sin(a), cz - half_width))
        v.append((cx + radius * math.cos(a), cy + radius * math.sin(a), cz + half_width))
    v.append((cx, cy, cz - half_width))
    v.append((cx, cy, cz + half_width))
    c0, c1 = 2 * segments, 2 * segments + 1
    t = []
    for k in range(segments):
        a0, a1 = 2 * k, 2 * k + 1
        b0, b1 = 2 * ((k + 1) % segments), 2 * ((k + 1) % segments) + 1
        t += [(a0, b0, b1), (a0, b1, a1), (c0, b0, a0), (c1, a1, b1)]
    return Mesh(v, t)


def merge_meshes(meshes):
    pos, idx, off = [], [], 0
    wheels = []
    for m in meshes:
        pos.append(m.positions)
        idx.append(m.indices + off)
        off += len(m.positions)
        wheels += m.wheels
    return Mesh(np.concatenate(pos), np.concatenate(idx), wheels)


def normalize_mesh(mesh, align_longest=True):
    """mesh.ts:271-318 for a y-up mesh: longest horizontal extent along +x, length 1, centred in x/z,
    lowest point at y = 0. Wheels follow the same transform (presets.ts:369-379)."""
    p = mesh.positions.astype(np.float64)
    T = np.eye(4)
    lo, hi = p.min(0), p.max(0)
    if align_longest and (hi[2] - lo[2]) > (hi[0] - lo[0]) * 1.05:
        R = np.eye(4)
        R[:3, :3] = rot_y(90)
        p = p @ R[:3, :3].T
        T = R @ T
        lo, hi = p.min(0), p.max(0)
    length = max(hi[0] - lo[0], 1e-9)
    cx, cz = (hi[0] + lo[0]) / 2, (hi[2] + lo[2]) / 2
    S = np.diag([1 / length, 1 / length, 1 / length, 1.0])
    Tr = np.eye(4)
    Tr[:3, 3] = [-cx, -lo[1], -cz]
    M = S @ Tr
    p = (p - np.array([cx, lo[1], cz])) / length
    T = M @ T
    wheels = []
    for w in mesh.wheels:
        c = (T @ np.array(list(w["c"]) + [1.0]))[:3]
        a = T[:3, :3] @ np.asarray(w["axis"], dtype=np.float64)
        a /= np.linalg.norm(a)
        wheels.append(dict(c=list(c), axis=list(a), r=w["r"] / length, hw=w["hw"] / length))
    return Mesh(p, mesh.indices, wheels)


def placement_matrix(mesh, dims, mode="ground", length_frac=1 / 3, diam_frac=0.3, x_frac=0.34, yaw_deg=0.0,
                     pitch_deg=0.0, ride_cells=0.0, spanwise=False):
    """Mesh -> grid matrix (column-major 16 floats) and scale (mesh.ts:354-379): scale to
    length_frac nx ('ground') or to diam_frac ny over the mesh height ('center'), rotate about the
    mid-height centre by yaw about y, then pitch about z (positive pitch lowers +x), translate to
    x_frac nx, z = nz / 2 and, on the ground, put the lowest vertex at ride_cells (the lowest of every
    k-th vertex, k = max(1, V // 200000), as the reference samples large meshes). spanwise stretches
    z to span the tunnel (nz + 4 cells; the quasi-2D cylinder)."""
    nx, ny, nz = dims
    lo, hi = mesh.min, mesh.max
    height = hi[1] - lo[1]
    scale = length_frac * nx if mode == "ground" else diam_frac * ny / max(height, 1e-6)
    cy = (hi[1] + lo[1]) / 2
    R = np.eye(4)
    R[:3, :3] = rot_y(yaw_deg) @ rot_z(-pitch_deg)
    sz = scale
    if spanwise:
        sz = (nz + 4) / max(hi[2] - lo[2], 1e-6) if nz > 1 else scale
    S = np.diag([scale, scale, sz, 1.0])
    C = np.eye(4)
    C[1, 3] = -cy
    M = R @ S @ C
    step = max(1, len(mesh.positions) // 200000)
    pts = np.c_[mesh.positions[::step].astype(np.float64), np.ones(len(mesh.positions[::step]))] @ M.T
    min_y = pts[:, 1].min()
    zc = nz / 2 if nz > 1 else 0.5
    ty = -min_y + ride_cells if mode == "ground" else ny / 2 + 0.37
    T = np.eye(4)
    T[:3, 3] = [x_frac * nx, ty, zc]
    full = T @ M
    return full.T.reshape(-1).astype(np.float32), scale


def grid_wheels(mesh, matrix16):
    """Wheels in grid space (mesh.ts:321-333)."""
    M = np.asarray(matrix16, dtype=np.float64).reshape(4, 4).T
    lin = M[:3, :3]
    s = np.linalg.norm(lin[:, 0])
    out = []
    for w in mesh.wheels:
        c = (M @ np.array(list(w["c"]) + [1.0]))[:3]
        a = lin @ np.asarray(w["axis"], dtype=np.float64)
        scale_axis = np.linalg.norm(a)
        a /= scale_axis
        out.append(dict(c=list(c), axis=list(a), r=w["r"] * s, hw=w["hw"] * scale_axis))
    return out
