"""The step with the learned closure (k_step_mlp: features from the live Pi^neq, BF16 MFMA network,
retau) against the FP64 oracle's step with the same closure (oracle/lbm_oracle.cpp, or_set_mlp),
after 100 steps from a seeded perturbed state, on the tiny-grid cases with a body, a rotating wheel,
the moving belt with a contact patch and both sponges (d3_full), and rows of three waves
(d3_long_rows). The network is a seeded random one with smag_floor = 0, so its inputs decide tau
everywhere off the body: nu_t spans about 0 to 2e-2 against nu = 1e-2.

Tolerances (max relative population error): FP16 deviation storage 2e-3 against the oracle with
FP16 store rounding, as for the plain step; FP32 storage 5e-5 against the FP64 oracle (the plain
step's bar is 1e-5). The network's inputs and first hidden layer are rounded to BF16, so a value
whose FP32 (device) and FP64 (oracle) forms straddle a BF16 rounding midpoint rounds to neighbouring
BF16 numbers on the two sides: about one cell in a thousand per step, each moving that cell's nu_t by
a fraction of a percent (the FP32 invariants tr S^3 and det S lose digits to cancellation, so the
device's inputs can be off by up to about 1e-4 near such a midpoint).
test_learned_step_single_collision_differs_at_rounding_midpoints checks that picture after a single
collision: at most 0.5 % of the network cells differ by more than 2e-6 (none by 2e-5), and those
cells are ones whose inputs lie unusually close to a BF16 midpoint (the median of their smallest
distance to one is under a fifth of that of all network cells; measured about a tenth).
"""
import numpy as np
import pytest

from oracle import mlp_oracle as M
from tests.cases import SEED, lattice_for, make_case, perturbed_state

pytestmark = pytest.mark.gpu

TOL = {"f32": 5e-5, "f16": 2e-3}
ORACLE_STORAGE = {"f32": None, "f16": "f16"}


def _net(seed=17):
    from windtunnel_vis_amd import closure
    net = closure.Network.initial(seed=seed, nu_scale=4e-3, smag_floor=0.0)
    net.mu = np.zeros(8, np.float32)
    net.sigma = np.full(8, 3.0, np.float32)
    return net


def _pair(name, prec, net):
    from oracle.oracle import OracleSolver
    from windtunnel_vis_amd import SolverGPU, closure
    dims, params, flags, sdf, wheels = make_case(name)
    p = dict(params, precision=prec)
    g = SolverGPU(dims, p)
    assert g.precision == prec
    o = OracleSolver(dims, p, storage=ORACLE_STORAGE[prec])
    f0 = perturbed_state(lattice_for(dims), dims, params["u"], flags, seed=SEED).astype(np.float32)
    for s, f in ((g, f0), (o, f0.astype(np.float64))):
        s.setWheels(wheels)
        s.uploadFlags(flags)
        s.uploadSdf(sdf)
        s.uploadPopulations(f)
    closure.attach(g, net)
    wd = closure.wall_distance(g)
    assert np.array_equal(wd, np.rint(M.wall_distance(flags, dims) * 32) / 32)
    o.setMlp(net, wd, pair=g.getTuning(1))
    return g, o


@pytest.mark.parametrize("prec", ["f32", "f16"])
@pytest.mark.parametrize("name", ["d3_full", "d3_long_rows"])
def test_learned_step_matches_oracle_100_steps(name, prec):
    net = _net()
    g, o = _pair(name, prec, net)
    print(f"[learned step parity] {name} {prec} cells/lane {g.getTuning(1)}", flush=True)
    try:
        g.step(100)
        o.step(100)
        a, b = g.readPopulations().astype(np.float64), o.readPopulations()
        assert np.all(np.isfinite(a))
        err = np.abs(a - b) / np.abs(b)
        k = int(np.argmax(err))
        print("  max rel err %.3e at dir %d cell %d" % (err.max(), k // g.n, k % g.n), flush=True)
        assert err.max() < TOL[prec]
        # the closure changed the flow: the same run with Smagorinsky differs by far more than the tolerance
        from oracle.oracle import OracleSolver
        dims, params, flags, sdf, wheels = make_case(name)
        ref = OracleSolver(dims, dict(params, precision=prec), storage=ORACLE_STORAGE[prec])
        ref.setWheels(wheels)
        ref.uploadFlags(flags)
        ref.uploadSdf(sdf)
        ref.uploadPopulations(perturbed_state(lattice_for(dims), dims, params["u"], flags, seed=SEED).astype(np.float32)
                              .astype(np.float64))
        ref.step(100)
        plain = np.abs(ref.readPopulations() - b) / np.abs(b)
        print("  learned vs Smagorinsky (oracle): %.3e" % plain.max(), flush=True)
        assert plain.max() > 5 * TOL[prec]
    finally:
        g.destroy()
        o.destroy()


@pytest.mark.parametrize("name", ["d3_full", "d3_long_rows"])
def test_learned_step_single_collision_differs_at_rounding_midpoints(name):
    net = _net()
    g, o = _pair(name, "f32", net)
    try:
        g.step(1)
        o.step(1)
        a, b = g.readPopulations().astype(np.float64), o.readPopulations()
        err = (np.abs(a - b) / np.abs(b)).max(axis=0)  # per cell
        margin = o.mlpMargins()
        net_cells = np.isfinite(margin)
        off = np.flatnonzero(err > 2e-6)
        m_all = float(np.median(margin[net_cells]))
        m_off = float(np.median(margin[off])) if len(off) else 0.0
        print(f"  {name}: {len(off)} of {int(net_cells.sum())} network cells differ by > 2e-6 (max {err.max():.2e}); "
              f"median BF16 margin {m_off:.2e} there against {m_all:.2e} over all network cells", flush=True)
        # isolated cells, small differences
        assert len(off) <= 0.005 * net_cells.sum() and err.max() < 2e-5
        # and they are cells whose inputs sit close to a BF16 midpoint: their margins are far below
        # the typical margin (FP32 differences that come from elsewhere would not pick such cells)
        assert m_off < 0.2 * m_all, [(int(c), float(err[c]), float(margin[c])) for c in off][:8]
    finally:
        g.destroy()
        o.destroy()


def test_learned_step_stats_match_oracle():
    """Force and statistics of the learned run (FP32 storage) against the oracle's."""
    net = _net(23)
    g, o = _pair("d3_full", "f32", net)
    try:
        for chunk in (0, 50, 50):
            sg, so = g.stepAndSample(chunk), o.stepAndSample(chunk)
            rg, ro = g.sampleRaw(), o.sampleRaw()
            scale = max(1.0, np.abs(ro[:3]).max())
            assert np.all(np.abs(rg[:3] - ro[:3]) <= 1e-4 * scale), (rg, ro)
            assert abs(rg[3] - ro[3]) <= 1e-5 * max(ro[3], 1e-6) + 1e-9
            assert sg.unstable == so.unstable and sg.step == so.step
    finally:
        g.destroy()
        o.destroy()
