"""CPU restatement (numpy, FP64) of the derived-field pass and the field export — test
infrastructure only; the product path never imports this module.

  * derived     — derivedWGSL (src/solver/wgslLbm.ts:527-592): central-difference velocity gradient
                  with clamped neighbour reads, vorticity, Q = (|W|^2 - |S|^2) / 2 faded out within
                  two cells of a solid (1 cell: off, 2 cells: halved; axis neighbours only, solid
                  tests false outside the domain), turbulence intensity from the EMA statistics.
  * derived_loops — the same pass as explicit per-cell loops (pins the vectorised form on tiny grids).
  * export_sample — SolverGPU.exportFields sampling (src/solver/SolverGPU.ts:374-425).
  * half        — IEEE binary16 rounding (the rgba16float stores).
  * sample / solid_at / streamlines / pcg_hash / particles_update — the tracer compute passes
    (src/render/tracers.ts:3-200): linear clamp-to-edge sampling, streamlines, particle tracers.
Pinned by closed forms (rigid rotations about each axis, pure shear, pure strain), the clamped-face
one-sided difference, the fade pattern around a single solid cell and the TI identity; the
sampler by linear fields (reproduced exactly inside, constant beyond the last texel centre); the
streamline and particle steps by uniform and rigid-rotation flows; the hash against an independent
JavaScript evaluation with the reference's u32 arithmetic.
"""
import math

import numpy as np


def half(x):
    """Round to IEEE half (nearest even), returned as float64."""
    return np.asarray(x, dtype=np.float64).astype(np.float16).astype(np.float64)


def _shift(a, axis, d):
    """a at index clamp(i + d) along `axis`."""
    n = a.shape[axis]
    idx = np.clip(np.arange(n) + d, 0, n - 1)
    return np.take(a, idx, axis=axis)


def _solid_shift(s, axis, d):
    """solid at i + d along `axis`, False outside the domain."""
    out = np.zeros_like(s)
    n = s.shape[axis]
    src = [slice(None)] * s.ndim
    dst = [slice(None)] * s.ndim
    if d > 0:
        src[axis] = slice(d, n)
        dst[axis] = slice(0, max(n - d, 0))
    else:
        src[axis] = slice(0, max(n + d, 0))
        dst[axis] = slice(-d, n)
    out[tuple(dst)] = s[tuple(src)]
    return out


def derived(vel4, solid, dims, meanA, meanB, uRef):
    """vel4 (n, 4) = (ux, uy, uz, rho) per cell, solid (n,) bool, meanA (n, 4), meanB (n,).
    Returns (vort (n, 4) = wx, wy, wz, Q; stat (n, 4) = TI, solid, 0, 0; mean (n, 4)), unrounded."""
    nx, ny, nz = dims
    dim = 3 if nz > 1 else 2
    u = np.asarray(vel4, dtype=np.float64).reshape(nz, ny, nx, 4)[..., :3]
    s = np.asarray(solid, dtype=bool).reshape(nz, ny, nx)
    # g[..., j, i] = d u_i / d x_j ; array axes (z, y, x) -> x_j axis 2 - j
    g = np.zeros((nz, ny, nx, 3, 3))
    for j in range(dim):
        ax = 2 - j
        g[..., j, :] = 0.5 * (_shift(u, ax, 1) - _shift(u, ax, -1))
    wx = g[..., 1, 2] - g[..., 2, 1]
    wy = g[..., 2, 0] - g[..., 0, 2]
    wz = g[..., 0, 1] - g[..., 1, 0]
    S = 0.5 * (g + np.swapaxes(g, -1, -2))
    W = 0.5 * (g - np.swapaxes(g, -1, -2))
    S2 = (S * S).sum(axis=(-1, -2))
    W2 = (W * W).sum(axis=(-1, -2))
    near = np.zeros((nz, ny, nx))
    for d, v in ((1, 1.0), (2, 0.5)):
        hit = np.zeros_like(s)
        for a in range(dim):
            ax = 2 - a
            hit |= _solid_shift(s, ax, d) | _solid_shift(s, ax, -d)
        near = np.where(hit, np.maximum(near, v), near)
    q = 0.5 * (W2 - S2) * (1.0 - near)
    vort = np.stack([wx, wy, wz, q], axis=-1)
    vort[s] = 0.0
    m = np.asarray(meanA, dtype=np.float64).reshape(-1, 4)
    k = np.maximum(np.asarray(meanB, dtype=np.float64).reshape(-1) - (m[:, :3] ** 2).sum(axis=1), 0.0)
    ti = np.sqrt(k / 3.0) / max(uRef, 1e-6)
    sf = s.reshape(-1)
    stat = np.zeros((len(sf), 4))
    stat[:, 0] = np.where(sf, 0.0, ti)
    stat[:, 1] = sf.astype(np.float64)
    return vort.reshape(-1, 4), stat, m.copy()


def derived_loops(vel4, solid, dims, meanA, meanB, uRef):
    """Per-cell loop form of `derived` (tiny grids only)."""
    nx, ny, nz = dims
    dim = 3 if nz > 1 else 2
    u = np.asarray(vel4, dtype=np.float64).reshape(-1, 4)
    s = np.asarray(solid, dtype=bool).reshape(-1)
    n = nx * ny * nz

    def at(x, y, z):
        x = min(max(x, 0), nx - 1); y = min(max(y, 0), ny - 1); z = min(max(z, 0), nz - 1)
        return u[x + nx * (y + ny * z), :3]

    def solid_at(x, y, z):
        if x < 0 or y < 0 or z < 0 or x >= nx or y >= ny or z >= nz:
            return False
        return bool(s[x + nx * (y + ny * z)])

    vort = np.zeros((n, 4))
    stat = np.zeros((n, 4))
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                i = x + nx * (y + ny * z)
                if not s[i]:
                    J = np.zeros((3, 3))  # J[i][j] = d u_i / d x_j
                    e = np.eye(3, dtype=int)
                    for j in range(dim):
                        J[:, j] = 0.5 * (at(x + e[j][0], y + e[j][1], z + e[j][2]) - at(x - e[j][0], y - e[j][1], z - e[j][2]))
                    w = (J[2, 1] - J[1, 2], J[0, 2] - J[2, 0], J[1, 0] - J[0, 1])
                    S = 0.5 * (J + J.T)
                    W = 0.5 * (J - J.T)
                    near = 0.0
                    for d in (1, 2):
                        for a in range(dim):
                            o = [0, 0, 0]
                            o[a] = d
                            if solid_at(x + o[0], y + o[1], z + o[2]) or solid_at(x - o[0], y - o[1], z - o[2]):
                                near = max(near, 1.0 if d == 1 else 0.5)
                    vort[i] = (w[0], w[1], w[2], 0.5 * ((W * W).sum() - (S * S).sum()) * (1.0 - near))
                m = np.asarray(meanA[i], dtype=np.float64)
                k = max(float(meanB[i]) - float((m[:3] ** 2).sum()), 0.0)
                stat[i] = (0.0 if s[i] else np.sqrt(k / 3.0) / max(uRef, 1e-6), 1.0 if s[i] else 0.0, 0.0, 0.0)
    return vort, stat


def export_dims(dims, stride):
    return tuple(-(-d // stride) for d in dims)


def export_sample(field4, dims, stride):
    """Rows of field4 (n, 4) at the sampled cells min(g * stride, n - 1), x fastest."""
    nx, ny, nz = dims
    ox, oy, oz = export_dims(dims, stride)
    gx = np.minimum(np.arange(ox) * stride, nx - 1)
    gy = np.minimum(np.arange(oy) * stride, ny - 1)
    gz = np.minimum(np.arange(oz) * stride, nz - 1)
    idx = (gx[None, None, :] + nx * (gy[None, :, None] + ny * gz[:, None, None])).reshape(-1)
    return np.asarray(field4).reshape(-1, 4)[idx]


# ---- probes, streamlines, particle tracers (render/tracers.ts) ----------------------------------
def sample(field4, dims, p):
    """Linear-filtered, clamp-to-edge sample of a (n, 4) cell field at points p (k, 3) in cell units
    (texel centres at i + 0.5; the single layer in 2D)."""
    nx, ny, nz = dims
    f = np.asarray(field4, dtype=np.float64).reshape(nz, ny, nx, 4)
    p = np.atleast_2d(np.asarray(p, dtype=np.float64))
    t = p - 0.5
    if nz == 1:
        t[:, 2] = 0.0
    out = np.zeros((len(p), 4))
    lo = np.floor(t)
    a = t - lo
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = (a[:, 0] if dx else 1 - a[:, 0]) * (a[:, 1] if dy else 1 - a[:, 1]) * (a[:, 2] if dz else 1 - a[:, 2])
                ix = np.clip(lo[:, 0] + dx, 0, nx - 1).astype(int)
                iy = np.clip(lo[:, 1] + dy, 0, ny - 1).astype(int)
                iz = np.clip(lo[:, 2] + dz, 0, nz - 1).astype(int)
                out += w[:, None] * f[iz, iy, ix]
    return out


def solid_at(solid, dims, p):
    nx, ny, nz = dims
    c = np.floor(np.atleast_2d(p)).astype(int)
    if nz == 1:
        c[:, 2] = 0
    inside = (c >= 0).all(1) & (c[:, 0] < nx) & (c[:, 1] < ny) & (c[:, 2] < nz)
    s = np.asarray(solid, dtype=bool).reshape(nz, ny, nx)
    out = np.zeros(len(c), bool)
    cc = c[inside]
    out[inside] = s[cc[:, 2], cc[:, 1], cc[:, 0]]
    return out


def _vel(field4, dims, p):
    v = sample(field4, dims, p)[:, :3]
    if dims[2] == 1:
        v[:, 2] = 0.0
    return v


def streamlines(field4, solid, dims, rake, seeds, per_line, points, h, U):
    """STREAMLINES_WGSL (tracers.ts:152-200): (seeds, points, 4)."""
    nx, ny, nz = dims
    per = max(per_line, 1)
    out = np.zeros((seeds, points, 4))
    a, b, d = (np.asarray(rake[k][:3], dtype=np.float64) for k in ("a", "b", "d"))
    for s in range(seeds):
        l = (s // per) - 0.5 * (seeds // per - 1)
        t = ((s % per) + 0.5) / per
        p = a + (b - a) * t + d * l
        if nz == 1:
            p[2] = 0.5
        alive = True
        for k in range(points):
            v = _vel(field4, dims, p)[0]
            sp = float(np.linalg.norm(v))
            outside = (p[0] < 0 or p[1] < 0 or p[0] >= nx or p[1] >= ny) if nz == 1 else (p.min() < 0 or p[0] >= nx or p[1] >= ny or p[2] >= nz)
            if alive and (outside or solid_at(solid, dims, p)[0] or sp < max(U, 1e-5) * 1e-3):
                alive = False
            out[s, k] = (*p, sp if alive else -1.0)
            if not alive:
                continue
            vm = _vel(field4, dims, p + v / sp * 0.5 * h)[0]
            lm = float(np.linalg.norm(vm))
            if lm < 1e-9:
                alive = False
                continue
            p = p + vm / lm * h
    return out


M32 = 0xFFFFFFFF


def pcg_hash(v):
    """PCG-RXS-M-XS 32-bit hash (Jarzynski & Olano, JCGT 2020), on Python ints."""
    s = (v * 747796405 + 2891336453) & M32
    w = (((s >> ((s >> 28) + 4)) ^ s) * 277803737) & M32
    return ((w >> 22) ^ w) & M32


class _Rng:
    def __init__(self, s):
        self.s = s

    def next(self):
        self.s = pcg_hash(self.s)
        return float(np.float32(self.s) / np.float32(4294967295.0))  # f32(u32) / f32 divisor, in f32


def _spawn(r, o, dims):
    nx, ny, nz = dims
    if o["emitter"] == 0:
        x = 1.0 + r.next() * 2.0
        y = 0.5 + r.next() * (ny - 1)
        z = 0.5 if nz == 1 else 0.5 + r.next() * (nz - 1)
        return np.array([x, y, z])
    lines, noz = max(o["lines"], 1), max(o["nozzles"], 1)
    k = min(int(r.next() * (noz * lines)), noz * lines - 1)
    t = ((k % noz) + 0.5) / noz
    l = (k // noz) - 0.5 * (lines - 1)
    a, b, d = (np.asarray(o["rake"][q][:3], dtype=np.float64) for q in ("a", "b", "d"))
    p = a + (b - a) * t + d * l
    p = p + (np.array([r.next(), r.next(), r.next()]) - 0.5) * 0.35
    if nz == 1:
        p[2] = 0.5
    return p


def particles_update(state, hist, head, field4, solid, dims, o, steps, seed, fill, advance):
    """PARTICLES_WGSL (tracers.ts:3-106) for all particles; updates state (count, 4) and hist
    (trail, count, 4) in place."""
    nx, ny, nz = dims
    for i in range(o["count"]):
        r = _Rng(pcg_hash(i ^ pcg_hash(seed)))
        s = state[i].copy()
        if fill:
            if o["emitter"] == 0:
                p = None
                for _ in range(8):
                    q = np.array([1.0 + r.next() * (nx - 2), 0.5 + r.next() * (ny - 1), 0.5 if nz == 1 else 0.5 + r.next() * (nz - 1)])
                    if not solid_at(solid, dims, q)[0]:
                        p = q
                        break
                if p is None:
                    p = _spawn(r, o, dims)
                s = np.array([*p, o["trail"]])
            else:
                p = _spawn(r, o, dims)
                s = np.array([*p, -math.floor(r.next() * o["max_age"])])
            state[i] = s
            hist[:, i] = [s[0], s[1], s[2], 0.0]
            continue
        if not advance:
            continue
        if s[3] < 0:
            s[3] += 1
            if s[3] >= 0:
                s = np.array([*_spawn(r, o, dims), 0.0])
        else:
            p = s[:3].copy()
            disp = float(np.linalg.norm(_vel(field4, dims, p)[0])) * steps
            nsub = int(min(max(math.ceil(disp / 0.6), 1), 24))
            h = steps / nsub
            for _ in range(nsub):
                a = _vel(field4, dims, p)[0]
                bb = _vel(field4, dims, p + a * 0.5 * h)[0]
                p = p + bb * h
            s = np.array([*p, s[3] + 1])
            outside = (p[0] < 0 or p[1] < 0 or p[0] >= nx - 1 or p[1] >= ny) if nz == 1 else \
                (p.min() < 0 or p[0] >= nx - 1 or p[1] >= ny or p[2] >= nz)
            if outside or solid_at(solid, dims, p)[0] or s[3] > o["max_age"] * 4:
                s = np.array([*_spawn(r, o, dims), 0.0])
        state[i] = s
        hist[head, i] = [s[0], s[1], s[2], 0.0]
