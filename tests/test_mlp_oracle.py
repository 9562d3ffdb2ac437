"""Pins of the learned-closure restatement (oracle/mlp_oracle.py) on the CPU: closed forms,
invariance, finite differences and an independent distance transform."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def _rot(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def _s6(S):
    return np.array([S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]])


def test_features_are_rotation_invariant_and_match_matrix_powers():
    rng = np.random.default_rng(3)
    for _ in range(20):
        A = rng.normal(size=(3, 3)) * 1e-3
        S = 0.5 * (A + A.T)
        R = _rot(rng)
        S2 = R @ S @ R.T
        x1 = M.raw_features(_s6(S)[None], 1.5, 0.05, 6e-4)
        x2 = M.raw_features(_s6(S2)[None], 1.5, 0.05, 6e-4)
        assert np.allclose(x1, x2, rtol=1e-9, atol=1e-9)
        # the invariants against matrix algebra (tr S^3 via Cayley-Hamilton in the restatement)
        r = M.S_REF
        assert np.isclose(x1[0, 4], M.slog(np.trace(S @ S @ S), r ** 3), rtol=1e-9)
        assert np.isclose(x1[0, 2], M.slog(np.linalg.det(S), r ** 3), rtol=1e-9)
        assert np.isclose(x1[0, 3], M.slog(np.trace(S @ S), r * r), rtol=1e-12)
        assert np.isclose(x1[0, 5], M.slog(np.sqrt(2 * np.sum(S * S)), r), rtol=1e-12)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    p = rng.normal(size=M.NPARAM) * 0.3
    x = rng.normal(size=(50, 8))
    t = np.abs(rng.normal(size=50))
    loss, g = M.loss_and_grad(p, x, t)
    for k in list(range(0, M.NPARAM, 37)) + [M.NPARAM - 1]:
        e = np.zeros(M.NPARAM)
        e[k] = 1e-6
        fd = (M.loss_and_grad(p + e, x, t)[0] - M.loss_and_grad(p - e, x, t)[0]) / 2e-6
        assert abs(fd - g[k]) < 1e-6 * max(1.0, abs(g[k])), k


def test_adam_first_step_is_signed_learning_rate():
    g = np.array([0.3, -2.0, 1e-3])
    p, m, v = M.adam(np.zeros(3), np.zeros(3), np.zeros(3), g, 1, lr=0.01)
    # bias-corrected first step: -lr * g / (|g| + eps)
    assert np.allclose(p, -0.01 * g / (np.abs(g) + 1e-8))


def test_wall_distance_matches_exact_distance_transform():
    from scipy.ndimage import distance_transform_edt
    rng = np.random.default_rng(5)
    dims = (14, 11, 9)
    flags = np.zeros(dims[::-1], np.uint32)
    flags[3:6, 4:8, 5:9] = 1
    flags[rng.random(flags.shape) < 0.01] = 1 | (1 << 8)
    d = M.wall_distance(flags.reshape(-1), dims)
    ref = np.minimum(distance_transform_edt(flags == 0), 4.0).reshape(-1)
    assert np.allclose(d, ref)


def test_coarse_samples_linear_shear():
    # u = (a y, 0, 0) on the fine grid: block variance of u_x is a^2 / 4 (tau_xx), the filtered
    # strain has S_xy = a / 2 (block means advance by 2 a per coarse cell), so tau_dev : S = 0 and
    # the target is 0; the inputs are those of S = [[0, a/2, 0], [a/2, 0, 0], [0, 0, 0]].
    a = 1e-3
    dims = (8, 12, 8)
    z, y, x = np.meshgrid(np.arange(8), np.arange(12), np.arange(8), indexing="ij")
    u = np.stack([a * (y + 0.5), 0 * y, 0 * y], axis=-1).reshape(-1, 3)
    solid = np.zeros(8 * 12 * 8, bool)
    wd = np.full(4 * 6 * 4, 4.0)
    feats, tgt, idx = M.coarse_samples(u, solid, dims, wd, 6e-4)
    assert len(idx) == 2 * 4 * 2  # interior coarse cells
    assert np.allclose(tgt, 0.0)
    S = np.zeros((1, 6))
    S[0, 3] = a  # S_xy = d(mean u_x)/dY / 2 with d/dY = 2 a
    ub = a * (2 * (idx // 4 % 6) + 1.0)
    ref = M.raw_features(np.repeat(S, len(idx), 0), np.full(len(idx), 4.0), ub, 6e-4)
    assert np.allclose(feats, ref)



def test_coarse_samples_one_sided_beside_solid():
    # The same shear with two coarse blocks holding a solid cell, at coarse (1, 2, 1) and (1, 4, 1):
    # the cells above and below them take one-sided differences, exact for a linear field, so their
    # strain is the same a / 2; cell (1, 3, 1) has both y neighbours bad and is left out.
    a = 1e-3
    dims = (8, 16, 8)
    z, y, x = np.meshgrid(np.arange(8), np.arange(16), np.arange(8), indexing="ij")
    u = np.stack([a * (y + 0.5), 0 * y, 0 * y], axis=-1).reshape(-1, 3)
    solid = np.zeros((8, 16, 8), bool)
    solid[2, 4, 2] = True  # fine cell of coarse block (1, 2, 1)
    solid[3, 9, 3] = True  # fine cell of coarse block (1, 4, 1)
    u[solid.reshape(-1)] = 0.0  # solid cells carry no velocity: a central difference across them is wrong
    wd = np.full(4 * 8 * 4, 4.0)
    feats, tgt, idx = M.coarse_samples(u, solid.reshape(-1), dims, wd, 6e-4)
    cx, cy = 4, 8
    cells = set((int(i) % cx, int(i) // cx % cy, int(i) // (cx * cy)) for i in idx)
    interior = set((x, y, z) for x in (1, 2) for y in range(1, 7) for z in (1, 2))
    assert cells == interior - {(1, 2, 1), (1, 4, 1), (1, 3, 1)}
    assert (1, 1, 1) in cells and (1, 5, 1) in cells  # one-sided along y
    assert np.allclose(tgt, 0.0)
    S = np.zeros((1, 6))
    S[0, 3] = a
    ub = a * (2 * (idx // cx % cy) + 1.0)
    ref = M.raw_features(np.repeat(S, len(idx), 0), np.full(len(idx), 4.0), ub, 6e-4)
    assert np.allclose(feats, ref)

def test_bf16_forward_close_to_fp64():
    rng = np.random.default_rng(6)
    p = rng.normal(size=M.NPARAM) * 0.3
    x = rng.normal(size=(200, 8))
    a = M.forward(p, x, 1e-3)[0]
    b = M.forward_bf16(p, x, 1e-3)
    assert np.max(np.abs(a - b) / (np.abs(a) + 1e-4)) < 0.05


def _linear_target(G):
    """Closed form of the training target for a fine field u = G x (G[i, k] = du_i / dx_k): the 2x2x2
    block variance of each coordinate is 1/4, so tau = G G^T / 4; block means advance by 2 G per coarse
    cell, so the filtered strain (coarse units) is S = G + G^T; nu_t = max(0, -tau_dev : S / (2 S : S))
    (uniform over the 3x3x3 neighbourhood)."""
    tau = G @ G.T / 4.0
    S = G + G.T
    tdev = tau - np.trace(tau) / 3.0 * np.eye(3)
    return max(0.0, -np.sum(tdev * S) / (2.0 * np.sum(S * S))), S


@pytest.mark.parametrize("seed", range(6))
def test_coarse_samples_linear_field_nonzero_target(seed):
    """A linear fine field u = u0 + G x with random G: the target has the closed form above, which is
    non-zero for about half of the draws (a flipped sign of the numerator sends those to 0 and the
    others to a positive value; a factor 2 in tau, S or the denominator scales it)."""
    rng = np.random.default_rng(100 + seed)
    G = rng.normal(size=(3, 3)) * 1e-3
    dims = (10, 12, 8)
    z, y, x = np.meshgrid(np.arange(8) + 0.5, np.arange(12) + 0.5, np.arange(10) + 0.5, indexing="ij")
    X = np.stack([x, y, z], axis=-1).reshape(-1, 3)
    u = np.array([0.05, 0.01, -0.02]) + X @ G.T
    solid = np.zeros(len(u), bool)
    wd = np.full(5 * 6 * 4, 4.0)
    feats, tgt, idx = M.coarse_samples(u, solid, dims, wd, 6e-4)
    assert len(idx) == 3 * 4 * 2
    want, S = _linear_target(G)
    assert np.allclose(tgt, want, rtol=1e-9, atol=1e-18)
    S6 = np.array([S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]])
    cx, cy = 5, 6
    c = np.stack([idx % cx, idx // cx % cy, idx // (cx * cy)], axis=1) * 2.0 + 1.0  # block centres, fine units
    ub = np.linalg.norm(np.array([0.05, 0.01, -0.02]) + c @ G.T, axis=1)
    ref = M.raw_features(np.repeat(S6[None], len(idx), 0), np.full(len(idx), 4.0), ub, 6e-4)
    assert np.allclose(feats, ref, rtol=1e-9, atol=1e-12)


def test_coarse_samples_nonzero_target_example():
    # u_x = a (x + y), a < 0: tau_xx = a^2 / 2, S_xx = 2a, S_xy = a, so nu_t = -a / 18 > 0; a > 0 gives 0
    for a, want in ((-1e-3, 1e-3 / 18), (1e-3, 0.0)):
        G = np.zeros((3, 3))
        G[0, 0] = G[0, 1] = a
        assert _linear_target(G)[0] == pytest.approx(want, rel=1e-12, abs=1e-20)
        dims = (8, 8, 8)
        z, y, x = np.meshgrid(np.arange(8) + 0.5, np.arange(8) + 0.5, np.arange(8) + 0.5, indexing="ij")
        u = np.stack([0.05 + a * (x + y), 0 * x, 0 * x], axis=-1).reshape(-1, 3)
        _, tgt, _ = M.coarse_samples(u, np.zeros(512, bool), dims, np.full(64, 4.0), 6e-4)
        assert np.allclose(tgt, want, rtol=1e-9, atol=1e-20)


def test_closure_restatements_agree():
    """The C++ oracle's closure network and features equal the numpy restatement (both FP64, the same
    BF16 roundings)."""
    from oracle import oracle as O
    from windtunnel_vis_amd import closure
    rng = np.random.default_rng(12)
    net = closure.Network.initial(seed=3, nu_scale=4e-3, smag_floor=0.0)
    net.mu = rng.normal(size=8).astype(np.float32)
    net.sigma = np.full(8, 3.0, np.float32)
    raw = rng.normal(size=(3000, 8)) * 3
    a = O.mlp_infer(net, raw)
    b = M.forward_bf16(net.params.astype(np.float64), (raw - net.mu) / net.sigma, net.nu_scale)
    assert np.allclose(a, b, rtol=1e-12, atol=0)
    S = rng.normal(size=(500, 6)) * 1e-3
    wd = rng.uniform(0, 4, 500)
    um = rng.uniform(0, 0.1, 500)
    assert np.allclose(O.mlp_features(S, wd, um, 6e-4), M.raw_features(S, wd, um, 6e-4), rtol=1e-13, atol=1e-13)
