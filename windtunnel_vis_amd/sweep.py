"""Seeded yaw / ride-height sweep of uneven D3Q19 domains (the batched-sweep workload).

Each case: a scale s ~ U[1, 2.5] gives a grid round16(128 s, 64 s, 64 s) (at most 320x160x160); the
body is the procedural sedan or SUV scaled to length nx / 3, yawed by an integer angle in
[-15, 15] degrees, with a ride height in {2..6} cells. Parameters: u = 0.06, Cs = 0.1, nu = 6e-4,
FP16 deviation storage, moving belt with rolling wheels, Bouzidi walls, the app's sponge / wall
defaults. Random numbers come from splitmix64 with the first seed 20261203.
"""
from . import geometry as G

SEED = 20261203
MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * (1.0 / (1 << 53))

    def integer(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + self.next() % (hi - lo + 1)


def round16(v):
    return max(16, int(round(v / 16.0)) * 16)


def sweep_params(precision="f16"):
    return dict(u=0.06, nu=6e-4, cs=0.1, ground="moving", sides="freeslip", spongeNu=0.04, spongeStart=0.84,
                spongeIn=0.03, tauWall=0.53, interp=True, collision="regularized", emaAlpha=0.02, precision=precision)


def sweep_cases(count=24, seed=SEED):
    """List of dicts {dims, kind, yaw, ride, scale}."""
    rng = SplitMix64(seed)
    cases = []
    for _ in range(count):
        s = 1.0 + 1.5 * rng.uniform()
        dims = (min(round16(128 * s), 320), min(round16(64 * s), 160), min(round16(64 * s), 160))
        kind = "suv" if rng.integer(0, 1) else "sedan"
        yaw = rng.integer(-15, 15)
        ride = rng.integer(2, 6)
        cases.append(dict(dims=dims, kind=kind, yaw=int(yaw), ride=int(ride), scale=s))
    return cases


def case_body(case):
    """The case's body, scaled from the 256-cell layout to length nx / 3."""
    nx = case["dims"][0]
    k = (nx / 3.0) / 85.0
    if case["kind"] == "suv":
        return G.sedan(case["dims"], length=85 * k, width=34 * k, height=30 * k, corner=7 * k, wheel_r=7 * k,
                       wheel_w=6 * k, ride=case["ride"], yaw_deg=case["yaw"])
    return G.sedan(case["dims"], length=85 * k, width=32 * k, height=22 * k, corner=6 * k, wheel_r=6 * k,
                   wheel_w=5 * k, ride=case["ride"], yaw_deg=case["yaw"])


def make_solvers(cases, precision="f16"):
    """One voxelized solver per case; returns [(solver, body, info)]."""
    from .solver import SolverGPU
    out = []
    for c in cases:
        s = SolverGPU(c["dims"], sweep_params(precision))
        body = case_body(c)
        s.setWheels(body.wheels)
        info = s.voxelizeSdf(body.prims())
        out.append((s, body, info))
    return out
