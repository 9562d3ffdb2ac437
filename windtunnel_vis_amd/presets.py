"""Procedural vehicle presets as closed triangle meshes for the mesh voxelizer (src/voxelize/presets.ts).

Bodies are lofted superellipse cross-sections along x (monotone cubic profiles for the roof, floor,
belt line and half width); wheels are closed cylinders along z, wings extruded NACA 4-digit
sections, plates boxes. Every part is closed and outward oriented on its own; the voxelizer's
winding-number rule resolves their union. `build_preset(id)` returns the mesh normalised to
length 1 (longest horizontal extent along +x, lowest point at y = 0, centred in x and z) with its
wheels in the same space, ready for `placement_matrix` and `SolverGPU.voxelizeMesh`.

The control points of the presets are the reference's: they define the shapes. The tessellation
is this module's own (the reference tessellates wheels, boxes, the sphere and the wing extrusion
with three.js, which is not available here), with the same segment counts.
"""
import math

import numpy as np

from .geometry import Mesh, box_mesh, merge_meshes, normalize_mesh, uv_sphere

PRESETS = [
    dict(id="sedan", name="Sedan", lengthM=4.8, placement="ground"),
    dict(id="sports", name="Sports car", lengthM=4.5, placement="ground"),
    dict(id="suv", name="SUV / van", lengthM=4.9, placement="ground"),
    dict(id="truck", name="Truck + trailer", lengthM=16.5, placement="ground"),
    dict(id="f1", name="Open-wheel racer", lengthM=5.5, placement="ground"),
    dict(id="ahmed25", name="Ahmed body 25° (reference)", lengthM=1.044, placement="ground"),
    dict(id="ahmed35", name="Ahmed body 35° (reference)", lengthM=1.044, placement="ground"),
    dict(id="sphere", name="Sphere (validation)", lengthM=0.2, placement="center", validation=True),
    dict(id="cylinder", name="Cylinder (validation)", lengthM=0.1, placement="center", spanwise=True, validation=True),
]


def monotone_cubic(pts):
    """Fritsch-Carlson monotone piecewise-cubic Hermite interpolant through (x, y) control points,
    constant outside them (presets.ts:34-61)."""
    xs = [float(p[0]) for p in pts]
    ys = [float(p[1]) for p in pts]
    n = len(xs)
    d = [(ys[i + 1] - ys[i]) / (xs[i + 1] - xs[i]) for i in range(n - 1)]
    m = [0.0] * n
    m[0], m[n - 1] = d[0], d[n - 2]
    for i in range(1, n - 1):
        m[i] = 0.0 if d[i - 1] * d[i] <= 0 else (d[i - 1] + d[i]) / 2
    for i in range(n - 1):
        if d[i] == 0:
            m[i] = m[i + 1] = 0.0
            continue
        a, b = m[i] / d[i], m[i + 1] / d[i]
        s = a * a + b * b
        if s > 9:
            t = 3 / math.sqrt(s)
            m[i], m[i + 1] = t * a * d[i], t * b * d[i]

    def f(x):
        if x <= xs[0]:
            return ys[0]
        if x >= xs[-1]:
            return ys[-1]
        i = 0
        while x > xs[i + 1]:
            i += 1
        h = xs[i + 1] - xs[i]
        t = (x - xs[i]) / h
        t2, t3 = t * t, t * t * t
        return ((2 * t3 - 3 * t2 + 1) * ys[i] + (t3 - 2 * t2 + t) * h * m[i] + (-2 * t3 + 3 * t2) * ys[i + 1]
                + (t3 - t2) * h * m[i + 1])
    return f


def loft(x0, x1, top, bot, belt, halfW, exp, tumble, expTop=None, stations=120, ring=64, zOffset=0.0):
    """Closed lofted body (presets.ts:81-133): `stations` + 1 rings of `ring` points clustered toward
    both ends (cosine spacing), each a superellipse of exponent `exp` (`expTop` above the belt) between
    the floor and the roof, half width `halfW`, the upper half narrowed by `tumble`; fan caps."""
    ftop, fbot, fbelt, fhw = monotone_cubic(top), monotone_cubic(bot), monotone_cubic(belt), monotone_cubic(halfW)
    e_top = exp if expTop is None else expTop

    def se(c, e):
        return math.copysign(abs(c) ** (2 / e), c)

    pos = []
    for s in range(stations + 1):
        u = s / stations
        t = 0.5 - 0.5 * math.cos(math.pi * u)
        x = x0 + (x1 - x0) * t
        yt, yb = ftop(x), fbot(x)
        yc = min(max(fbelt(x), yb + 1e-4), yt - 1e-4)
        w = max(fhw(x), 1e-4)
        for r in range(ring):
            phi = 2 * math.pi * r / ring
            sy = math.sin(phi)
            upper = sy > 0
            e = e_top if upper else exp
            cz = se(math.cos(phi), e)
            cy = se(sy, e)
            y = yc + cy * ((yt - yc) if upper else (yc - yb))
            z = w * cz * ((1 - tumble * cy) if upper else 1.0)
            pos.append((x, y, z + zOffset))
    idx = []

    def rg(s, r):
        return s * ring + (r % ring)

    for s in range(stations):
        for r in range(ring):
            a, b, c, d = rg(s, r), rg(s, r + 1), rg(s + 1, r), rg(s + 1, r + 1)
            idx += [(a, c, b), (b, c, d)]
    for s in (0, stations):
        ctr = np.mean([pos[rg(s, r)] for r in range(ring)], axis=0)
        pos.append(tuple(ctr))
    c0, c1 = len(pos) - 2, len(pos) - 1
    for r in range(ring):
        idx.append((c0, rg(0, r), rg(0, r + 1)))
        idx.append((c1, rg(stations, r + 1), rg(stations, r)))
    return Mesh(pos, idx)


def cylinder_z(radius, length, segments, center=(0.0, 0.0, 0.0)):
    """Closed cylinder of `segments` sides with its axis along z, outward oriented."""
    cx, cy, cz = center
    h = length / 2
    v = []
    for k in range(segments):
        a = 2 * math.pi * k / segments
        x, y = cx + radius * math.sin(a), cy + radius * math.cos(a)
        v += [(x, y, cz - h), (x, y, cz + h)]
    v += [(cx, cy, cz - h), (cx, cy, cz + h)]
    c0, c1 = 2 * segments, 2 * segments + 1
    t = []
    for k in range(segments):
        a0, a1 = 2 * k, 2 * k + 1
        b0, b1 = 2 * ((k + 1) % segments), 2 * ((k + 1) % segments) + 1
        t += [(a0, b1, b0), (a0, a1, b1), (c0, a0, b0), (c1, b1, a1)]
    return orient(Mesh(v, t))


def cylinder_y(radius, length, segments, center):
    """Closed cylinder with its axis along y (the Ahmed body's stilts)."""
    m = cylinder_z(radius, length, segments)
    p = m.positions.astype(np.float64)[:, [0, 2, 1]] + np.asarray(center, dtype=np.float64)
    return orient(Mesh(p, m.indices))


def signed_volume(mesh):
    """Volume enclosed by an outward-oriented closed mesh (negative when inward oriented)."""
    p = mesh.positions.astype(np.float64)
    a, b, c = p[mesh.indices[:, 0]], p[mesh.indices[:, 1]], p[mesh.indices[:, 2]]
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)


def orient(mesh):
    """The part with its triangles flipped if it encloses a negative volume (mesh.ts: mergeObject)."""
    if signed_volume(mesh) < 0:
        return Mesh(mesh.positions, mesh.indices[:, ::-1], mesh.wheels)
    return mesh


class _Builder:
    def __init__(self):
        self.parts = []
        self.wheels = []

    def add(self, mesh):
        self.parts.append(orient(mesh))

    def wheel(self, r, width, x, z, y=None):
        """Wheel of radius r and width `width` at (x, y = r, z), axis z, 40 sides (presets.ts:139-145)."""
        y = r if y is None else y
        self.wheels.append(dict(c=[x, y, z], axis=[0.0, 0.0, 1.0], r=r, hw=width / 2))
        self.parts.append(cylinder_z(r, width, 40, (x, y, z)))

    def plate(self, x0, x1, y0, y1, z0, z1):
        self.parts.append(box_mesh((x0, y0, z0), (x1, y1, z1)))

    def mesh(self):
        m = merge_meshes(self.parts)
        m.wheels = list(self.wheels)
        return m


def rounded_box(x0, x1, y0, y1, halfW, exp=8, zOffset=0.0):
    """presets.ts:147-158."""
    L = x1 - x0
    r = min(0.04 * L, (y1 - y0) * 0.1)
    return loft(x0, x1,
                top=[(x0, y0 + (y1 - y0) * 0.6), (x0 + r * 0.5, y1 - r * 0.3), (x0 + r * 2, y1), (x1 - r, y1), (x1, y1 - r * 0.5)],
                bot=[(x0, y0 + (y1 - y0) * 0.3), (x0 + r * 0.5, y0 + r * 0.2), (x0 + r * 2, y0), (x1, y0)],
                belt=[(x0, (y0 + y1) / 2), (x1, (y0 + y1) / 2)],
                halfW=[(x0, halfW * 0.8), (x0 + r * 1.5, halfW), (x1 - r, halfW), (x1, halfW * 0.92)],
                exp=exp, tumble=0.0, stations=60, ring=48, zOffset=zOffset)


def airfoil(chord, camber, camberPos, thick, aoaDeg, N=40):
    """NACA 4-digit section, quarter-chord centred and rotated by the angle of attack (presets.ts:161-189):
    upper surface from the leading to the trailing edge, then the lower surface back (2N points)."""
    up, lo = [], []
    for i in range(N + 1):
        b = i / N * math.pi
        x = 0.5 - 0.5 * math.cos(b)
        yt = 5 * thick * (0.2969 * math.sqrt(x) - 0.126 * x - 0.3516 * x * x + 0.2843 * x ** 3 - 0.1036 * x ** 4)
        yc = dyc = 0.0
        if camber > 0:
            if x < camberPos:
                yc = camber / camberPos ** 2 * (2 * camberPos * x - x * x)
                dyc = 2 * camber / camberPos ** 2 * (camberPos - x)
            else:
                yc = camber / (1 - camberPos) ** 2 * (1 - 2 * camberPos + 2 * camberPos * x - x * x)
                dyc = 2 * camber / (1 - camberPos) ** 2 * (camberPos - x)
        th = math.atan(dyc)
        up.append((x - yt * math.sin(th), yc + yt * math.cos(th)))
        lo.append((x + yt * math.sin(th), yc - yt * math.cos(th)))
    a = math.radians(aoaDeg)

    def tr(p):
        px, py = (p[0] - 0.25) * chord, p[1] * chord
        return (px * math.cos(a) + py * math.sin(a), -px * math.sin(a) + py * math.cos(a))

    return [tr(p) for p in up] + [tr(p) for p in lo[1:-1][::-1]]


def wing(x, y, chord, span, aoaDeg, camber=0.06):
    """Inverted (downforce) airfoil extruded along z over [-span/2, span/2], centred at (x, y)
    (presets.ts:192-198). The caps pair the upper and lower surface points into a strip."""
    pts = airfoil(chord, camber, 0.4, 0.12, aoaDeg)
    n = len(pts)  # 2N: upper 0..N, lower N+1..2N-1 (trailing to leading edge)
    N = n // 2
    v = []
    for z in (-span / 2, span / 2):
        for px, py in pts:
            v.append((x + px, y - py, z))
    t = []
    for k in range(n):  # side walls
        a0, a1 = k, (k + 1) % n
        t += [(a0, a1, n + a1), (a0, n + a1, n + k)]
    # caps: strip between the upper point i and the lower point at the same chord station
    low = lambda i: 0 if i == 0 else (N if i == N else n - i)  # index of the lower point at station i
    for i in range(N):
        u0, u1, l0, l1 = i, i + 1, low(i), low(i + 1)
        for q in ((u0, u1, l1), (u0, l1, l0)):
            if len(set(q)) < 3:
                continue
            t.append((q[2], q[1], q[0]))      # cap at -span/2: against the side walls' ring edges
            t.append((n + q[0], n + q[1], n + q[2]))
    return orient(Mesh(v, t))


def sedan():
    b = _Builder()
    b.add(loft(0, 1,
               top=[(0, 0.085), (0.015, 0.125), (0.06, 0.16), (0.27, 0.183), (0.31, 0.19), (0.44, 0.292), (0.5, 0.302), (0.63, 0.298),
                    (0.76, 0.255), (0.84, 0.21), (0.96, 0.205), (1, 0.165)],
               bot=[(0, 0.06), (0.03, 0.04), (0.1, 0.032), (0.9, 0.036), (0.97, 0.06), (1, 0.1)],
               belt=[(0, 0.07), (0.1, 0.13), (0.5, 0.15), (0.9, 0.16), (1, 0.14)],
               halfW=[(0, 0.1), (0.04, 0.165), (0.15, 0.19), (0.85, 0.19), (0.97, 0.175), (1, 0.14)],
               exp=3.2, expTop=2.6, tumble=0.32))
    for x in (0.165, 0.79):
        for z in (-0.158, 0.158):
            b.wheel(0.068, 0.048, x, z)
    return b.mesh()


def sports():
    b = _Builder()
    b.add(loft(0, 1,
               top=[(0, 0.06), (0.015, 0.09), (0.08, 0.12), (0.3, 0.15), (0.35, 0.162), (0.47, 0.24), (0.55, 0.252), (0.66, 0.235),
                    (0.85, 0.19), (0.96, 0.17), (1, 0.14)],
               bot=[(0, 0.045), (0.03, 0.028), (0.1, 0.024), (0.9, 0.026), (0.98, 0.06), (1, 0.08)],
               belt=[(0, 0.05), (0.12, 0.1), (0.5, 0.12), (0.85, 0.135), (1, 0.12)],
               halfW=[(0, 0.1), (0.05, 0.185), (0.2, 0.205), (0.75, 0.215), (0.95, 0.2), (1, 0.16)],
               exp=3.0, expTop=2.3, tumble=0.4))
    for x, r in ((0.17, 0.07), (0.79, 0.074)):
        for z in (-0.172, 0.172):
            b.wheel(r, 0.058, x, z)
    return b.mesh()


def suv():
    b = _Builder()
    b.add(loft(0, 1,
               top=[(0, 0.13), (0.012, 0.175), (0.05, 0.21), (0.24, 0.235), (0.29, 0.25), (0.4, 0.352), (0.46, 0.366), (0.93, 0.362),
                    (0.975, 0.34), (1, 0.25)],
               bot=[(0, 0.09), (0.03, 0.055), (0.1, 0.048), (0.9, 0.05), (0.98, 0.07), (1, 0.1)],
               belt=[(0, 0.12), (0.1, 0.19), (0.9, 0.21), (1, 0.2)],
               halfW=[(0, 0.13), (0.03, 0.175), (0.12, 0.192), (0.9, 0.192), (0.99, 0.18), (1, 0.16)],
               exp=4.5, expTop=3.2, tumble=0.18))
    for x in (0.17, 0.8):
        for z in (-0.158, 0.158):
            b.wheel(0.078, 0.052, x, z)
    return b.mesh()


def truck():
    b = _Builder()
    b.add(loft(0, 0.15,
               top=[(0, 0.1), (0.01, 0.17), (0.03, 0.195), (0.07, 0.205), (0.1, 0.228), (0.13, 0.238), (0.15, 0.236)],
               bot=[(0, 0.045), (0.02, 0.03), (0.15, 0.03)],
               belt=[(0, 0.1), (0.15, 0.12)],
               halfW=[(0, 0.066), (0.015, 0.074), (0.15, 0.076)],
               exp=7, expTop=5, tumble=0.05, stations=60, ring=56))
    b.plate(0.1, 0.2, 0.03, 0.055, -0.035, 0.035)
    b.add(rounded_box(0.17, 1.0, 0.075, 0.245, 0.076, 10))
    for z in (-0.058, 0.058):
        b.wheel(0.031, 0.02, 0.07, z)
    for x in (0.17, 0.23, 0.87, 0.92, 0.97):
        for z in (-0.058, 0.058):
            b.wheel(0.031, 0.024, x, z)
    return b.mesh()


def f1():
    b = _Builder()
    b.add(loft(0.03, 0.93,
               top=[(0.03, 0.045), (0.06, 0.06), (0.15, 0.078), (0.3, 0.1), (0.36, 0.108), (0.4, 0.172), (0.46, 0.17), (0.62, 0.12),
                    (0.8, 0.085), (0.93, 0.065)],
               bot=[(0.03, 0.04), (0.06, 0.035), (0.15, 0.04), (0.28, 0.018), (0.85, 0.018), (0.93, 0.03)],
               belt=[(0.03, 0.043), (0.15, 0.06), (0.35, 0.075), (0.6, 0.06), (0.93, 0.045)],
               halfW=[(0.03, 0.01), (0.08, 0.02), (0.2, 0.035), (0.3, 0.05), (0.4, 0.052), (0.6, 0.04), (0.85, 0.028), (0.93, 0.02)],
               exp=2.6, expTop=2.2, tumble=0.3))
    for z in (-0.085, 0.085):
        b.add(loft(0.33, 0.72,
                   top=[(0.33, 0.06), (0.36, 0.085), (0.45, 0.09), (0.6, 0.07), (0.72, 0.035)],
                   bot=[(0.33, 0.03), (0.36, 0.022), (0.72, 0.022)],
                   belt=[(0.33, 0.05), (0.72, 0.03)],
                   halfW=[(0.33, 0.02), (0.36, 0.042), (0.5, 0.04), (0.72, 0.012)],
                   exp=3, tumble=0.2, stations=50, ring=40, zOffset=z))
    b.plate(0.3, 0.86, 0.012, 0.02, -0.135, 0.135)
    b.add(wing(0.03, 0.028, 0.07, 0.36, -4))
    b.add(wing(0.085, 0.045, 0.04, 0.34, -18))
    for z in (-0.18, 0.18):
        b.plate(0.0, 0.11, 0.012, 0.065, z - 0.003, z + 0.003)
    b.add(wing(0.9, 0.16, 0.06, 0.18, -8))
    b.add(wing(0.945, 0.18, 0.035, 0.18, -25))
    for z in (-0.09, 0.09):
        b.plate(0.86, 0.98, 0.09, 0.2, z - 0.003, z + 0.003)
    b.plate(0.9, 0.93, 0.06, 0.16, -0.004, 0.004)
    for z in (-0.145, 0.145):
        b.wheel(0.06, 0.055, 0.17, z)
        b.wheel(0.062, 0.072, 0.77, z)
    return b.mesh()


def ahmed(slantDeg):
    """Ahmed reference body (Ahmed, Ramm & Faltin 1984): 1044 x 389 x 288 mm, front rounded with
    R = 100 mm, a 222 mm rear slant at `slantDeg`, 50 mm ground clearance on four 30 mm stilts
    (presets.ts:318-354)."""
    L, W, H, R, Ls, c = 1044.0, 389.0, 288.0, 100.0, 222.0, 50.0
    k = 1 / L
    tan = math.tan(math.radians(slantDeg))

    def rnd(x):
        return R - math.sqrt(max(R * R - (R - x) ** 2, 0.0)) if x < R else 0.0

    N = 80
    xs = []
    for i in range(N + 1):
        u = i / N
        xs.append(u / 0.3 * R * 1.2 if u < 0.3 else R * 1.2 + (u - 0.3) / 0.7 * (L - R * 1.2))
    top, bot, hw, belt = [], [], [], []
    for x in xs:
        r = rnd(x)
        t = c + H - r
        if x > L - Ls:
            t = min(t, c + H - (x - (L - Ls)) * tan)
        top.append((x * k, t * k))
        bot.append((x * k, (c + r) * k))
        hw.append((x * k, (W / 2 - r) * k))
        belt.append((x * k, (c + H / 2) * k))
    b = _Builder()
    b.add(loft(0, 1, top, bot, belt, hw, exp=14, tumble=0.0, stations=140, ring=72))
    for x in (163.0, 163.0 + 470.0):
        for z in (-(W / 2 - 60), W / 2 - 60):
            b.add(cylinder_y(15 * k, (c + 20) * k, 20, (x * k, (c + 20) / 2 * k, z * k)))
    return b.mesh()


def sphere():
    return uv_sphere(0.5, 96, 64)


def cylinder():
    """Spanwise validation cylinder: diameter 1, length 1 along z, 128 sides."""
    return cylinder_z(0.5, 1.0, 128)


BUILDERS = dict(sedan=sedan, sports=sports, suv=suv, truck=truck, f1=f1, sphere=sphere, cylinder=cylinder,
                ahmed25=lambda: ahmed(25), ahmed35=lambda: ahmed(35))


def build_preset(pid):
    """The preset `pid` normalised (presets.ts:361-380): length 1 along +x, ground at y = 0."""
    if pid not in BUILDERS:
        raise KeyError("unknown preset %r (have %s)" % (pid, ", ".join(BUILDERS)))
    return normalize_mesh(BUILDERS[pid](), align_longest=pid != "cylinder")
