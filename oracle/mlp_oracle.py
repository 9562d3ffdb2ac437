"""CPU restatement (numpy, FP64) of the learned-closure pieces — test infrastructure only.

The product path never imports this module; tests compare the HIP kernels against it:
  * raw_features        — csrc/mlp.hpp mlp_raw_features (six strain invariants, wall distance, local Re)
  * wall_distance       — csrc/kernels_mlp.hip k_wdist_* (nearest solid cell centre within 4 cells)
  * coarse_samples      — csrc/kernels_mlp.hip k_coarse_fields / k_coarse_samples (2x box filter of
                          a fine run, subgrid stress, least-squares eddy viscosity target)
  * forward / backward  — the 8-32-32-1 network (ReLU, softplus output) and its MSE gradient
  * forward_bf16        — the MFMA inference path's roundings (BF16 weights and activations)
  * adam                — the optimizer step of k_adam
There is no reference implementation of this path (the closure is new); each function is pinned by
tests against closed forms, finite differences or brute force.
"""
import numpy as np

S_REF = 1e-3
NPARAM = 1377
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 256, 288, 1312, 1344, 1376


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(32, 8), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(32, 32), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3], p[OFF_B3])


def slog(x, ref):
    return np.sign(x) * np.log1p(np.abs(x) / ref)


def raw_features(S, wd, umag, nu0):
    """S: (n, 6) = sxx, syy, szz, sxy, sxz, syz. Returns (n, 8)."""
    S = np.asarray(S, dtype=np.float64)
    sxx, syy, szz, sxy, sxz, syz = S.T
    i1 = sxx + syy + szz
    tr2 = sxx ** 2 + syy ** 2 + szz ** 2 + 2 * (sxy ** 2 + sxz ** 2 + syz ** 2)
    i2 = 0.5 * (i1 * i1 - tr2)
    i3 = sxx * (syy * szz - syz ** 2) - sxy * (sxy * szz - syz * sxz) + sxz * (sxy * syz - syy * sxz)
    tr3 = i1 ** 3 - 3 * i1 * i2 + 3 * i3
    m = np.sqrt(2 * tr2)
    r = S_REF
    wd = np.broadcast_to(np.asarray(wd, dtype=np.float64), i1.shape)
    um = np.broadcast_to(np.asarray(umag, dtype=np.float64), i1.shape)
    x = np.stack([slog(i1, r), slog(i2, r * r), slog(i3, r ** 3), slog(tr2, r * r), slog(tr3, r ** 3), slog(m, r),
                  wd, np.log1p(um * np.maximum(wd, 0.5) / nu0)], axis=1)
    return x


def softplus(z):
    return np.where(z > 20, z, np.log1p(np.exp(np.minimum(z, 20))))


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(p, xn, nu_scale):
    """xn: normalised inputs (n, 8). Returns (nu, cache)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    a1 = xn @ W1.T + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0)
    z = h2 @ w3 + b3
    return nu_scale * np.minimum(softplus(z), TARGET_CLIP), (xn, a1, h1, a2, h2, z)


TARGET_CLIP = 10.0


def loss_and_grad(p, xn, target):
    """Mean squared error of softplus(z) against target (nu_t / nu_scale, clipped at TARGET_CLIP
    by the caller); gradient w.r.t. p."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    _, (x, a1, h1, a2, h2, z) = forward(p, xn, 1.0)
    pred = softplus(z)
    n = len(z)
    r = pred - target
    loss = np.mean(r * r)
    dz = 2.0 * r / n * sigmoid(z)
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = dz @ h2
    g[OFF_B3] = dz.sum()
    d2 = np.outer(dz, w3) * (a2 > 0)
    g[OFF_W2:OFF_B2] = (d2.T @ h1).reshape(-1)
    g[OFF_B2:OFF_W3] = d2.sum(0)
    d1 = (d2 @ W2) * (a1 > 0)
    g[OFF_W1:OFF_B1] = (d1.T @ x).reshape(-1)
    g[OFF_B1:OFF_W2] = d1.sum(0)
    return loss, g


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(p, xn, nu_scale):
    """The MFMA path: BF16 inputs, weights and hidden activations; FP32-class accumulation."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    h1 = np.maximum(bf16(xn) @ bf16(W1).T + b1, 0)
    h2 = np.maximum(bf16(h1) @ bf16(W2).T + b2, 0)
    z = bf16(h2) @ bf16(w3) + b3
    return nu_scale * np.minimum(softplus(z), TARGET_CLIP)


def adam(p, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def wall_distance(flags, dims, radius=4):
    """Distance (cells) from every cell centre to the nearest solid cell centre, capped at `radius`;
    0 on solid cells. Brute force over the (2r+1)^3 cube."""
    nx, ny, nz = dims
    solid = np.asarray(flags).reshape(nz, ny, nx) != 0
    best = np.full((nz, ny, nx), float(radius * radius))
    for dz in range(-radius, radius + 1):
        for dy in range(-radius, radius + 1):
            for dx in range(-radius, radius + 1):
                d2 = dx * dx + dy * dy + dz * dz
                if d2 > radius * radius:
                    continue
                sh = np.zeros_like(solid)
                zs = slice(max(dz, 0), nz + min(dz, 0)); zd = slice(max(-dz, 0), nz + min(-dz, 0))
                ys = slice(max(dy, 0), ny + min(dy, 0)); yd = slice(max(-dy, 0), ny + min(-dy, 0))
                xs = slice(max(dx, 0), nx + min(dx, 0)); xd = slice(max(-dx, 0), nx + min(-dx, 0))
                sh[zd, yd, xd] = solid[zs, ys, xs]
                best = np.where(sh & (d2 < best), d2, best)
    return np.sqrt(best).reshape(-1)


def coarse_samples(u_fine, solid_fine, dims_fine, wd_coarse, nu0):
    """Training samples from a fine velocity field (n_f, 3): 2x box filter, subgrid stress
    tau = <u u> - <u><u>, filtered strain by central differences on the coarse grid (one-sided along
    an axis with one solid-holding neighbour block), target nu_t = max(0, sum(-tau_dev : S) /
    sum(2 S : S)) over the 3x3x3 neighbourhood (cells without a strain contribute 0). A coarse cell
    is used when its 8 fine cells are fluid, it is off the domain faces and along every axis at
    least one neighbour block is all fluid. Returns (features (m, 8), target (m,), index (m,))."""
    nx, ny, nz = dims_fine
    cx, cy, cz = nx // 2, ny // 2, nz // 2
    u = np.asarray(u_fine, dtype=np.float64).reshape(nz, ny, nx, 3)
    s = np.asarray(solid_fine).reshape(nz, ny, nx)

    def block(a):
        return a.reshape(cz, 2, cy, 2, cx, 2, *a.shape[3:]).mean(axis=(1, 3, 5))

    ub = block(u)
    uu = block(u[..., :, None] * u[..., None, :])
    bad = block(s.astype(np.float64)) > 0
    tau = uu - ub[..., :, None] * ub[..., None, :]
    g = np.zeros((cz, cy, cx, 3, 3))  # g[..., i, j] = d u_j / d x_i
    valid = ~bad
    for ax, i in ((2, 0), (1, 1), (0, 2)):
        # central difference; one-sided where one neighbour block holds a solid cell
        up, dn = ~np.roll(bad, -1, axis=ax), ~np.roll(bad, 1, axis=ax)
        u_up = np.where(up[..., None], np.roll(ub, -1, axis=ax), ub)
        u_dn = np.where(dn[..., None], np.roll(ub, 1, axis=ax), ub)
        h = np.where(up & dn, 0.5, 1.0)[..., None]
        g[..., i, :] = h * (u_up - u_dn)
        valid &= up | dn
    valid[:, :, 0] = valid[:, :, -1] = False
    valid[:, 0, :] = valid[:, -1, :] = False
    valid[0, :, :] = valid[-1, :, :] = False
    S = 0.5 * (g + np.swapaxes(g, -1, -2))
    tr = np.trace(tau, axis1=-2, axis2=-1)
    tdev = tau - tr[..., None, None] / 3.0 * np.eye(3)
    ss = np.where(valid, 2.0 * np.einsum("...ij,...ij->...", S, S), 0.0)
    num = np.where(valid, -np.einsum("...ij,...ij->...", tdev, S), 0.0)
    # least-squares eddy viscosity of the 3x3x3 neighbourhood (valid cells are interior, so no wrap)
    N3 = np.zeros_like(num)
    D3 = np.zeros_like(ss)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                N3 += np.roll(num, (-dz, -dy, -dx), axis=(0, 1, 2))
                D3 += np.roll(ss, (-dz, -dy, -dx), axis=(0, 1, 2))
    nut = np.maximum(0.0, N3 / (D3 + 1e-30))
    idx = np.flatnonzero(valid.reshape(-1))
    Sv = S.reshape(-1, 3, 3)[idx]
    S6 = np.stack([Sv[:, 0, 0], Sv[:, 1, 1], Sv[:, 2, 2], Sv[:, 0, 1], Sv[:, 0, 2], Sv[:, 1, 2]], axis=1)
    um = np.linalg.norm(ub.reshape(-1, 3)[idx], axis=1)
    wd = np.asarray(wd_coarse, dtype=np.float64).reshape(-1)[idx]
    return raw_features(S6, wd, um, nu0), nut.reshape(-1)[idx], idx
