"""Pins of the oracle's step with the learned closure (oracle/lbm_oracle.cpp, or_set_mlp) on the CPU.

The closure replaces the Smagorinsky tau of the collision (wgslLbm.ts:214) by
tau = t0 + 3 max(nu_net, floor nu_smag), nu_net = network(features of S = -3 Pi^neq / (2 rho tau_smag),
wall distance, local Re). Pinned here by:
  - a uniform state in a periodic box: one step is one collision, so the relaxation of Pi^neq
    measures tau; it must equal the closed form built from the numpy restatement of the features
    and the network (a wrong strain scale, sign or index, a wrong viscosity or guard fails it);
  - a constant network: the plain collision with nu + c;
  - a network with ~0 output and the full guard: Smagorinsky, also on a body with wheels and belt.
"""
import math

import numpy as np
import pytest

from oracle import mlp_oracle as M
from oracle.oracle import OracleSolver
from tests.cases import box_body_3d, make_case, perturbed_state
from tests.test_oracle_pins import _moments, _uniform_state
from windtunnel_vis_amd.solver import D3Q19


def _net(seed, nu_scale, floor, mu=0.0, sigma=3.0):
    from windtunnel_vis_amd import closure
    net = closure.Network.initial(seed=seed, nu_scale=nu_scale, smag_floor=floor)
    net.mu = np.full(8, mu, np.float32)
    net.sigma = np.full(8, sigma, np.float32)
    return net


@pytest.mark.parametrize("seed,floor,cs", [(1, 0.0, 0.17), (2, 0.0, 0.0), (3, 0.5, 0.17), (4, 1.0, 0.3)])
def test_closure_tau_in_uniform_box(seed, floor, cs):
    lat = D3Q19
    dims = (8, 9, 6)
    n = int(np.prod(dims))
    nu = 0.005
    tau0 = 3 * nu + 0.5
    pi = np.array([[2e-3, 0.0, -4e-4], [0.0, -1e-3, 0.0], [-4e-4, 0.0, -1e-3]])
    u = np.array([0.05, 0.0, 0.03])
    rho = 1.02
    f0 = _uniform_state(lat, n, rho, u, pi)
    net = _net(seed, 2e-3, floor)
    o = OracleSolver(dims, dict(u=0.05, nu=nu, cs=cs, sides="periodic", spongeNu=0.0), periodic_x=True)
    o.uploadPopulations(f0)
    o.setMlp(net, np.full(n, 4.0), pair=2)
    o.step(1)
    y = (np.arange(n) // dims[0]) % dims[1]
    inner = (y > 0) & (y < dims[1] - 1)
    _, _, P0 = _moments(lat, f0[:, inner])
    _, _, P1 = _moments(lat, o.readPopulations()[:, inner])
    # closed form of tau
    P = P0[:, :, 0]
    piN = math.sqrt(float((P ** 2).sum()))
    tau_s = 0.5 * (tau0 + math.sqrt(tau0 ** 2 + 18 * math.sqrt(2) * cs * cs * piN / rho))
    S = -3.0 * P / (2.0 * rho * tau_s)
    S6 = np.array([[S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]]])
    x = M.raw_features(S6, 4.0, np.linalg.norm(u), nu)
    nu_net = float(M.forward_bf16(net.params.astype(np.float64), (x - net.mu) / net.sigma, net.nu_scale)[0])
    assert 1e-5 < nu_net < 2e-2  # the network's output matters
    tau = tau0 + 3 * max(nu_net, floor * (tau_s - tau0) / 3)
    assert np.allclose(P1, (1 - 1 / tau) * P0, atol=1e-13, rtol=0)
    # and differs measurably from plain Smagorinsky
    assert abs(tau - tau_s) > 1e-4


def test_constant_network_is_added_viscosity():
    # no body, free-slip ground, Cs = 0: every cell runs the network, whose output is the constant c
    dims = (16, 10, 8)
    n = int(np.prod(dims))
    _, params, _, _, _ = make_case("d3_full")
    p = dict(params, ground="freeslip", cs=0.0, nu=0.01)
    c = 2e-3
    net = _net(0, 1e-3, 0.0)
    net.params[:] = 0
    net.params[M.OFF_B3] = np.log(np.expm1(c / 1e-3))
    flags = np.zeros(n, np.uint32)
    f0 = perturbed_state(D3Q19, dims, p["u"], flags)
    a = OracleSolver(dims, p)
    b = OracleSolver(dims, dict(p, nu=0.01 + c))
    for s in (a, b):
        s.uploadFlags(flags)
        s.uploadPopulations(f0)
    a.setMlp(net, np.full(n, 4.0))
    a.step(20)
    b.step(20)
    fa, fb = a.readPopulations(), b.readPopulations()
    # tau0 + 3 c against 3 (nu + c) + 1/2 and log(expm1) round trip: round-off over 20 steps
    assert np.max(np.abs(fa - fb)) < 1e-10


def test_zero_network_with_full_guard_is_smagorinsky_on_a_body():
    dims = (24, 12, 10)
    flags, sdf, wheels = box_body_3d(dims)
    _, params, _, _, _ = make_case("d3_full")
    net = _net(0, 1e-3, 1.0)
    net.params[M.OFF_B3] = -40.0
    f0 = perturbed_state(D3Q19, dims, params["u"], flags)
    a, b = OracleSolver(dims, params), OracleSolver(dims, params)
    for s in (a, b):
        s.setWheels(wheels)
        s.uploadFlags(flags)
        s.uploadSdf(sdf)
        s.uploadPopulations(f0)
    a.setMlp(net, M.wall_distance(flags, dims))
    a.step(30)
    b.step(30)
    assert np.max(np.abs(a.readPopulations() - b.readPopulations())) < 1e-13


def test_random_network_changes_the_flow_only_off_the_body():
    # cells next to the body (pair groups holding a wall or solid cell) keep Smagorinsky: after one
    # step from the same state, populations there equal the plain run's; bulk cells differ
    dims = (24, 12, 10)
    flags, sdf, wheels = box_body_3d(dims)
    _, params, _, _, _ = make_case("d3_full")
    net = _net(5, 4e-3, 0.0)
    f0 = perturbed_state(D3Q19, dims, params["u"], flags)
    a, b = OracleSolver(dims, params), OracleSolver(dims, params)
    for s in (a, b):
        s.setWheels(wheels)
        s.uploadFlags(flags)
        s.uploadSdf(sdf)
        s.uploadPopulations(f0)
    a.setMlp(net, M.wall_distance(flags, dims))
    a.step(1)
    b.step(1)
    d = np.max(np.abs(a.readPopulations() - b.readPopulations()), axis=0)
    # the plain-pull wall classification, by brute force
    nx, ny, nz = dims
    fl = flags.reshape(nz, ny, nx) != 0
    wall = np.zeros_like(fl)
    for i in range(1, 19):
        cx, cy, cz = D3Q19.c[i]
        for z in range(nz):
            for yy in range(ny):
                for x in range(nx):
                    sx, sy, sz = min(x - cx, nx - 1), yy - cy, z - cz
                    if cx > 0 and sx < 0:
                        continue
                    if not (0 <= sy < ny and 0 <= sz < nz):
                        continue
                    if fl[sz, sy, sx] and not fl[z, yy, x]:
                        wall[z, yy, x] = True
    bad = (wall | fl).reshape(-1)
    grp = bad.reshape(-1, 2).any(axis=1).repeat(2)
    assert np.all(d[grp] == 0.0)
    assert np.mean(d[~grp] > 1e-9) > 0.9
