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
