"""Seeded tiny-grid setups shared by the parity tests: every boundary branch of the pull step
(inlet, outlet clamp, free-slip mirrors, y-z edge, periodic z, ground bounce-back / moving belt with
a tyre contact patch, halfway and Bouzidi walls, rotating wheel with mass fix), both collision
operators, Smagorinsky, both sponges, FP32 / FP16 / BF16 storage."""
import numpy as np

from windtunnel_vis_amd.solver import D2Q9, D3Q19

SEED = 20261203


def feq(lat, rho, u):
    """Equilibrium populations, shape (q, n), FP64. rho: (n,), u: (3, n)."""
    cu = lat.c.T.astype(np.float64).T @ u  # (q, n)
    usq = (u * u).sum(axis=0)
    return lat.w[:, None] * rho[None, :] * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq[None, :])


def perturbed_state(lat, dims, U, flags, seed=SEED, amp_rho=2e-3, amp_u=0.02):
    """feq(1 + d_rho, (U,0,0) + d_u) with seeded noise on fluid cells, rest state on solids."""
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    rho = 1.0 + amp_rho * rng.uniform(-1, 1, n)
    u = np.zeros((3, n))
    u[0] = U
    u[0] += amp_u * rng.uniform(-1, 1, n)
    u[1] += amp_u * rng.uniform(-1, 1, n)
    if lat.dim == 3:
        u[2] += amp_u * rng.uniform(-1, 1, n)
    solid = flags != 0
    rho[solid] = 1.0
    u[:, solid] = 0.0
    return feq(lat, rho, u)


def box_body_3d(dims, with_wheel=True):
    """A small block sitting on a wheel near the floor plus a floating block, with distances."""
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    # floating rounded-ish block
    cx, cy, cz = 0.45 * nx, 0.55 * ny, 0.5 * nz
    hx, hy, hz = 2.6, 1.8, 2.2
    qx, qy, qz = np.abs(x - cx) - hx, np.abs(y - cy) - hy, np.abs(z - cz) - hz
    d_box = np.sqrt(np.maximum(qx, 0) ** 2 + np.maximum(qy, 0) ** 2 + np.maximum(qz, 0) ** 2) + np.minimum(np.maximum(qx, np.maximum(qy, qz)), 0)
    d = d_box
    wheels = []
    if with_wheel:
        # wheel: cylinder along z touching the floor, radius 2.2, half width 1.5
        wc = (0.7 * nx, 2.2, 0.5 * nz)
        rr = np.hypot(x - wc[0], y - wc[1]) - 2.2
        ww = np.abs(z - wc[2]) - 1.5
        d_wh = np.minimum(np.maximum(rr, ww), 0) + np.hypot(np.maximum(rr, 0), np.maximum(ww, 0))
        d = np.minimum(d, d_wh)
        wheels = [dict(c=list(wc), axis=[0, 0, 1], r=2.2, hw=1.5)]
        # a non-wheel block on the floor upstream (contact patch holds the belt)
        qx2, qy2, qz2 = np.abs(x - 0.2 * nx) - 1.0, np.abs(y - 1.0) - 1.0, np.abs(z - 0.3 * nz) - 1.0
        d_b2 = np.sqrt(np.maximum(qx2, 0) ** 2 + np.maximum(qy2, 0) ** 2 + np.maximum(qz2, 0) ** 2) + np.minimum(np.maximum(qx2, np.maximum(qy2, qz2)), 0)
        d = np.minimum(d, d_b2)
    flags = (d < 0).astype(np.uint32)
    if with_wheel:
        in_wheel = (np.hypot(x - wc[0], y - wc[1]) <= 2.2 + 0.75) & (np.abs(z - wc[2]) <= 1.5 + 0.6) & (d < 0)
        flags[in_wheel] = 1 | (1 << 8)
    sdf = np.where(np.abs(d) < 2.5, np.abs(d), -1.0).astype(np.float32)
    return flags.reshape(-1), sdf.reshape(-1), wheels


def cylinder_2d(dims, r=3.3):
    nx, ny, _ = dims
    y, x = np.meshgrid(np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    dd = np.hypot(x - 0.4 * nx, y - (ny / 2 + 0.37)) - r
    flags = (dd < 0).astype(np.uint32).reshape(-1)
    sdf = np.where(np.abs(dd) < 2.5, np.abs(dd), -1.0).astype(np.float32).reshape(-1)
    return flags, sdf


def make_case(name):
    """Returns (dims, params, flags, sdf, wheels)."""
    base3 = dict(u=0.06, nu=0.01, cs=0.1, ground="moving", sides="freeslip", spongeNu=0.04, spongeStart=0.8,
                 spongeIn=0.15, tauWall=0.53, interp=True, collision="regularized", emaAlpha=0.1)
    base2 = dict(u=0.06, nu=0.02, cs=0.1, ground="noslip", sides="freeslip", spongeNu=0.03, spongeStart=0.8,
                 spongeIn=0.1, tauWall=0.5, interp=True, collision="regularized", emaAlpha=0.1)
    if name == "d3_full":
        dims = (24, 12, 10)
        f, s, w = box_body_3d(dims)
        return dims, dict(base3), f, s, w
    if name == "d3_periodic_bgk":
        dims = (24, 12, 10)
        f, s, w = box_body_3d(dims)
        p = dict(base3, sides="periodic", collision="bgk", ground="noslip")
        return dims, p, f, s, w
    if name == "d3_freeslip_halfway":
        dims = (20, 10, 8)
        f, s, w = box_body_3d(dims, with_wheel=False)
        p = dict(base3, ground="freeslip", interp=False, spongeIn=0.0, cs=0.0)
        return dims, p, f, s, w
    if name == "d3_odd_f32":
        dims = (17, 9, 7)  # odd n: the last cell has no pair partner
        f, s, w = box_body_3d(dims)
        return dims, dict(base3), f, s, w
    if name == "d3_long_rows":
        dims = (384, 6, 5)  # three waves per row: the middle one holds neither row end
        f, s, w = box_body_3d(dims)
        return dims, dict(base3), f, s, w
    if name == "d2_long_rows":
        dims = (300, 10, 1)
        f, s = cylinder_2d(dims)
        return dims, dict(base2), f, s, []
    if name == "d2_cyl":
        dims = (40, 20, 1)
        f, s = cylinder_2d(dims)
        return dims, dict(base2), f, s, []
    if name == "d2_moving_bgk":
        dims = (40, 20, 1)
        f, s = cylinder_2d(dims)
        p = dict(base2, ground="moving", collision="bgk", interp=False)
        return dims, p, f, s, []
    raise KeyError(name)


def lattice_for(dims):
    return D3Q19 if dims[2] > 1 else D2Q9
