"""Learned closure on the GPU against the numpy restatement (oracle/mlp_oracle.py):
wall distance (bit-exact), training records from a fine run (FP32 vs FP64), the training
kernels' gradient and Adam steps, the MFMA inference path, and the closure inside the step."""
import numpy as np
import pytest

from oracle import mlp_oracle as M
from tests.cases import SEED, box_body_3d, lattice_for, make_case, perturbed_state

pytestmark = pytest.mark.gpu


def _solver(dims, params, flags, sdf, wheels, seed=SEED):
    from windtunnel_vis_amd import SolverGPU
    s = SolverGPU(dims, params)
    s.setWheels(wheels)
    s.uploadFlags(flags)
    s.uploadSdf(sdf)
    s.uploadPopulations(perturbed_state(lattice_for(dims), dims, params["u"], flags, seed=seed).astype(np.float32))
    return s


def test_wall_distance_bit_exact():
    from windtunnel_vis_amd import closure
    dims = (30, 14, 12)
    flags, sdf, wheels = box_body_3d(dims)
    _, params, _, _, _ = make_case("d3_full")
    s = _solver(dims, dict(params, precision="f16"), flags, sdf, wheels)
    got = closure.wall_distance(s)
    ref = np.rint(M.wall_distance(flags, dims) * 32) / 32
    assert np.array_equal(got, ref)
    s.destroy()


def _fine_coarse(prec):
    _, params, _, _, _ = make_case("d3_full")
    fd, cd = (40, 24, 24), (20, 12, 12)
    ff, fs, fw = box_body_3d(fd)
    cf, cs, cw = box_body_3d(cd)
    fine = _solver(fd, dict(params, precision=prec, nu=0.02), ff, fs, fw)
    coarse = _solver(cd, dict(params, precision=prec, nu=0.01), cf, cs, cw)
    fine.step(30)
    return fine, coarse, ff, cf, fd, cd


@pytest.mark.parametrize("prec", ["f32", "f16"])
def test_dataset_matches_restatement(prec):
    from windtunnel_vis_amd import closure
    fine, coarse, ff, cf, fd, cd = _fine_coarse(prec)
    ds = closure.Dataset(100000)
    added = ds.add(fine, coarse)
    rec = ds.read()
    assert added == len(rec) > 50
    idx = rec[:, 9].view(np.uint32)
    assert np.all(np.diff(idx.astype(np.int64)) > 0)  # cell order
    f = fine.readPopulations().astype(np.float64)
    lat = lattice_for(fd)
    rho = f.sum(0)
    u = (lat.c.T.astype(np.float64) @ f / rho).T
    wd = closure.wall_distance(coarse)
    feats, tgt, ridx = M.coarse_samples(u, ff != 0, fd, wd, 0.01)
    assert np.array_equal(idx, ridx.astype(np.uint32))
    assert np.allclose(rec[:, :8], feats, rtol=2e-4, atol=2e-4)
    scale = np.max(tgt) + 1e-12
    assert np.max(np.abs(rec[:, 8] - tgt)) < 2e-3 * scale
    ds.destroy()
    fine.destroy()
    coarse.destroy()


def _synthetic(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 8)) * 2 + 1
    t = 1e-3 * np.log1p(np.exp(0.5 * x[:, 0] - 0.3 * x[:, 5] + 0.2 * x[:, 6] * x[:, 7]))
    rec = np.zeros((n, 10), np.float32)
    rec[:, :8] = x
    rec[:, 8] = t
    rec[:, 9] = np.arange(n, dtype=np.uint32).view(np.float32)
    return rec


def _net(seed=7):
    from windtunnel_vis_amd import closure
    net = closure.Network.initial(seed=seed, nu_scale=1e-3)
    net.mu = np.full(8, 1.0, np.float32)
    net.sigma = np.full(8, 2.0, np.float32)
    return net


def test_gradient_matches_restatement():
    from windtunnel_vis_amd import closure
    rec = _synthetic(1000, 1)
    ds = closure.Dataset(2000)
    ds.append(rec)
    net = _net()
    first, count = 100, 700
    loss, g = closure.gradient(ds, net, first, count)
    r = rec[first:first + count].astype(np.float64)
    xn = (r[:, :8] - net.mu) / net.sigma
    lref, gref = M.loss_and_grad(net.params.astype(np.float64), xn, np.minimum(r[:, 8] / net.nu_scale, M.TARGET_CLIP))
    assert abs(loss - lref) < 1e-5 * lref
    assert np.max(np.abs(g - gref)) < 1e-4 * np.max(np.abs(gref))
    ds.destroy()


def test_full_batch_adam_steps_match_restatement():
    from windtunnel_vis_amd import closure
    rec = _synthetic(512, 2)
    ds = closure.Dataset(512)
    ds.append(rec)
    net = _net(8)
    p0 = net.params.astype(np.float64).copy()
    losses = closure.train(ds, net, epochs=3, batch=512, lr=1e-3)
    r = rec.astype(np.float64)
    xn = (r[:, :8] - net.mu) / net.sigma
    t = np.minimum(r[:, 8] / net.nu_scale, M.TARGET_CLIP)
    p, m, v = p0, np.zeros_like(p0), np.zeros_like(p0)
    ref_losses = []
    for k in range(1, 4):
        l, g = M.loss_and_grad(p, xn, t)
        ref_losses.append(l)
        p, m, v = M.adam(p, m, v, g, k, 1e-3)
    assert np.allclose(losses, ref_losses, rtol=1e-5)
    assert np.max(np.abs(net.params - p)) < 1e-5
    ds.destroy()


def test_training_reduces_loss():
    from windtunnel_vis_amd import closure
    rec = _synthetic(60000, 3)
    ds = closure.Dataset(60000)
    ds.append(rec)
    net = _net(9)
    losses = closure.train(ds, net, epochs=8, batch=1024, lr=3e-3)
    print("losses", losses, flush=True)
    assert losses[-1] < 0.2 * losses[0]
    ds.destroy()


def test_mfma_inference_matches_bf16_restatement():
    from windtunnel_vis_amd import closure
    rng = np.random.default_rng(11)
    net = _net(10)
    net.params = (rng.normal(size=M.NPARAM) * 0.4).astype(np.float32)
    raw = (rng.normal(size=(1000, 8)) * 2 + 1).astype(np.float32)  # not a multiple of 128 cells
    nu = closure.infer(net, raw)
    xn = (raw.astype(np.float64) - net.mu) / net.sigma
    ref = M.forward_bf16(net.params.astype(np.float64), xn, net.nu_scale)
    full = M.forward(net.params.astype(np.float64), xn, net.nu_scale)[0]
    err = np.abs(nu - ref) / (np.abs(ref) + 1e-5)
    print("max rel err vs bf16 restatement %.2e, vs fp64 %.2e" % (err.max(), (np.abs(nu - full) / (np.abs(full) + 1e-5)).max()),
          flush=True)
    assert np.median(err) < 2e-3 and np.max(err) < 3e-2


def test_constant_network_equals_plain_collision_with_added_viscosity():
    # a network that outputs nu_t = c everywhere turns the MLP closure into the plain collision with
    # nu + c (no body: no wall cells; free-slip ground: no ground-wall floor; Cs = 0: the
    # Smagorinsky guard is 0)
    from windtunnel_vis_amd import SolverGPU, closure
    dims = (32, 16, 12)
    _, params, _, _, _ = make_case("d3_full")
    p = dict(params, ground="freeslip", precision="f32", cs=0.0, nu=0.01)
    flags = np.zeros(32 * 16 * 12, np.uint32)
    sdf = np.full(32 * 16 * 12, -1.0, np.float32)
    c = 2e-3
    net = closure.Network(nu_scale=1e-3, smag_floor=0.0)
    net.params[M.OFF_B3] = np.log(np.expm1(c / 1e-3))  # softplus(b3) * 1e-3 = c
    a = _solver(dims, p, flags, sdf, [])
    closure.attach(a, net)
    b = _solver(dims, dict(p, nu=0.01 + c), flags, sdf, [])
    a.step(20)
    b.step(20)
    fa, fb = a.readPopulations(), b.readPopulations()
    err = np.max(np.abs(fa - fb)) / np.max(np.abs(fb))
    print("constant-network vs plain: %.2e" % err, flush=True)
    assert err < 2e-5
    a.destroy()
    b.destroy()


def test_zero_network_with_full_guard_is_smagorinsky():
    # with smag_floor = 1 and a network that outputs ~0 the closure is Smagorinsky everywhere: the
    # whole learned path (features, MFMA, retau, ground-wall BGK rows, skipped and solid lanes) on
    # a geometry with a body, wheels and a moving belt reproduces the plain run
    from windtunnel_vis_amd import closure
    dims = (40, 16, 16)
    flags, sdf, wheels = box_body_3d(dims)
    _, params, _, _, _ = make_case("d3_full")
    p = dict(params, precision="f32")
    net = closure.Network(nu_scale=1e-3, smag_floor=1.0)
    net.params[M.OFF_B3] = -30.0
    a = _solver(dims, p, flags, sdf, wheels)
    closure.attach(a, net)
    b = _solver(dims, p, flags, sdf, wheels)
    a.step(50)
    b.step(50)
    fa, fb = a.readPopulations(), b.readPopulations()
    err = np.max(np.abs(fa - fb)) / np.max(np.abs(fb))
    print("zero network + guard vs Smagorinsky: %.2e" % err, flush=True)
    assert err < 1e-5
    a.destroy()
    b.destroy()
