"""Lattice <-> SI mapping of a run (src/analysis/units.ts).

The lattice uses dx = dt = 1 and rho0 = 1; a run shares only its Reynolds number with the real
flow. The simulated Reynolds number is the real one scaled by the factor that puts the top speed
at what the grid resolves (nuMin), so it grows with the speed inside the simulator.
"""
import math
from decimal import ROUND_HALF_UP, Decimal

AIR_RHO = 1.225   # kg/m^3
AIR_NU = 1.5e-5   # m^2/s at 20 C
CS = 1.0 / math.sqrt(3.0)
U_MAX_LATTICE = 0.12


def latticeSpeed(speed, vMax):
    """Inlet speed in cells/step: 0.03 at rest up to 0.08 at vMax (units.ts:50-53)."""
    s = min(max(speed / vMax, 0.0), 1.0)
    return 0.03 + (0.08 - 0.03) * s


def mapFlow(speed, lengthM, nuAir, Lcells, vMax, nuMin, reOverride=None, uScale=1.0, nuScale=1.0):
    """units.ts:62-74: returns reReal, reSim, U, nu, tau, mach, dx (m per cell), dt (s per step)."""
    reReal = speed * lengthM / nuAir
    U = latticeSpeed(speed, vMax) * uScale
    reMaxGrid = 0.08 * Lcells / nuMin
    reRealMax = vMax * lengthM / AIR_NU
    reSim = reOverride if reOverride and reOverride > 0 else reReal * (reMaxGrid / reRealMax)
    nu = U * Lcells / reSim
    nu = max(nu * nuScale, nuMin * 0.5)
    reSim = U * Lcells / nu
    dx = lengthM / Lcells
    dt = (U / speed) * dx if speed > 0 else 0.0
    return dict(reReal=reReal, reSim=reSim, U=U, nu=nu, tau=3 * nu + 0.5, mach=U / CS, dx=dx, dt=dt)


def _to_fixed(x, digits):
    """Number.prototype.toFixed for x >= 0: the exact value rounded half up."""
    return str(Decimal(x).quantize(Decimal(1).scaleb(-digits), rounding=ROUND_HALF_UP))


def formatRe(re):
    """units.ts:76-82."""
    if not math.isfinite(re):
        return "—"
    if re >= 1e6:
        return _to_fixed(re / 1e6, 2) + "·10⁶"
    if re >= 1e4:
        return _to_fixed(re / 1e3, 0) + "k"
    if re >= 1e3:
        return _to_fixed(re / 1e3, 1) + "k"
    return _to_fixed(re, 0)
