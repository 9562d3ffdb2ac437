"""Result and field writers (src/analysis/export.ts): CSV tables and legacy-VTK field files.

The files are byte-for-byte what the reference writes for the same data: numbers in CSV cells are
rounded to 7 significant digits and printed the way JavaScript prints a Number (export.ts:19-23),
and the VTK file is the binary big-endian STRUCTURED_POINTS layout of export.ts:29-57.
"""
import math
from decimal import ROUND_HALF_UP, Decimal

import numpy as np


def js_number(v):
    """String(v) for a finite double v: the shortest round-trip digits, written in plain notation
    when the decimal exponent n (v = 0.d1..dk x 10^n) is in (-6, 21], else as d.ddde+-x
    (ECMAScript Number::toString)."""
    v = float(v)
    if v == 0.0:
        return "0"
    if v < 0:
        return "-" + js_number(-v)
    _, digits, exp = Decimal(repr(v)).as_tuple()
    d = "".join(map(str, digits)).rstrip("0")
    if not d:
        return "0"
    k = len(d)
    n = exp + len(digits)  # position of the decimal point relative to the first digit
    if k <= n <= 21:
        return d + "0" * (n - k)
    if 0 < n <= 21:
        return d[:n] + "." + d[n:]
    if -6 < n <= 0:
        return "0." + "0" * (-n) + d
    e = n - 1
    sign = "+" if e >= 0 else "-"
    return (d if k == 1 else d[0] + "." + d[1:]) + "e" + sign + str(abs(e))


def js_to_precision(v, p):
    """The Number `+v.toPrecision(p)`: v rounded to p significant digits, ties away from zero
    (toPrecision picks the larger candidate of an exact tie), read back as a double."""
    v = float(v)
    if v == 0.0:
        return 0.0
    x = Decimal(v)  # exact binary value
    e = x.adjusted()
    q = x.quantize(Decimal(1).scaleb(e - p + 1), rounding=ROUND_HALF_UP)
    return float(q)


def _cell(v):
    if isinstance(v, str):
        return v
    v = float(v)
    return js_number(js_to_precision(v, 7)) if math.isfinite(v) else ""


def csv(header, rows, comments=()):
    """export.ts:19-23: '# ' comment lines, the header, one line per row; numbers at 7 significant
    digits, non-finite numbers as empty cells. Returns the text (newline-terminated)."""
    lines = ["# %s" % c for c in comments] + [",".join(header)]
    for r in rows:
        lines.append(",".join(_cell(v) for v in r))
    return "\n".join(lines) + "\n"


def stride_for(n, max_points):
    """Export stride keeping at most ~max_points points (backend/webgpu.ts:133-137)."""
    return max(1, math.ceil(np.cbrt(n / max_points)))


def vtk(f, dx, U, L, rhoRef, title):
    """Legacy VTK (binary, big-endian) STRUCTURED_POINTS file of a FieldExport (export.ts:29-57):
    velocity and mean velocity / U, vorticity L / U, Cp = 2 (rho - rhoRef) / (3 U^2) of the
    instantaneous and mean density, Q L^2 / U^2; origin half a cell in, spacing dx * stride.
    Returns the file as bytes."""
    nx, ny, nz = f.dims
    n = nx * ny * nz
    h = dx * f.stride
    vel = np.asarray(f.vel, dtype=np.float32).reshape(n, 4).astype(np.float64)
    mean = np.asarray(f.mean, dtype=np.float32).reshape(n, 4).astype(np.float64)
    vort = np.asarray(f.vort, dtype=np.float32).reshape(n, 4).astype(np.float64)
    parts = []

    def text(s):
        parts.append(s.encode("utf-8"))

    def floats(a):
        parts.append(np.ascontiguousarray(a, dtype=">f4").tobytes())
        text("\n")

    def cp(rho):
        return (2.0 * (rho - rhoRef)) / (3.0 * U * U)

    text("# vtk DataFile Version 3.0\n%s\nBINARY\nDATASET STRUCTURED_POINTS\n" % title[:250])
    text("DIMENSIONS %d %d %d\nORIGIN %s %s %s\nSPACING %s %s %s\nPOINT_DATA %d\n"
         % (nx, ny, nz, js_number(0.5 * dx), js_number(0.5 * dx), js_number(0.5 * dx) if nz > 1 else "0",
            js_number(h), js_number(h), js_number(h), n))
    text("VECTORS velocity float\n")
    floats(vel[:, :3] / U)
    text("VECTORS mean_velocity float\n")
    floats(mean[:, :3] / U)
    text("VECTORS vorticity float\n")
    floats((vort[:, :3] * L) / U)
    text("SCALARS Cp float 1\nLOOKUP_TABLE default\n")
    floats(cp(vel[:, 3]))
    text("SCALARS mean_Cp float 1\nLOOKUP_TABLE default\n")
    floats(cp(mean[:, 3]))
    text("SCALARS Q_criterion float 1\nLOOKUP_TABLE default\n")
    floats((vort[:, 3] * L * L) / (U * U))
    return b"".join(parts)


def read_vtk(data):
    """Parses a file written by vtk(): returns (header dict, {name: array}) with vectors as (n, 3)."""
    pos = 0
    hdr = {}
    arrays = {}

    def line():
        nonlocal pos
        e = data.index(b"\n", pos)
        s = data[pos:e].decode("utf-8")
        pos = e + 1
        return s

    hdr["version"] = line()
    hdr["title"] = line()
    hdr["format"] = line()
    hdr["dataset"] = line()
    hdr["dims"] = tuple(int(v) for v in line().split()[1:])
    hdr["origin"] = line().split()[1:]
    hdr["spacing"] = line().split()[1:]
    n = int(line().split()[1])
    while pos < len(data):
        kind = line().split()
        comps = 3 if kind[0] == "VECTORS" else 1
        if kind[0] == "SCALARS":
            line()  # LOOKUP_TABLE default
        a = np.frombuffer(data, dtype=">f4", count=n * comps, offset=pos).astype(np.float32)
        pos += 4 * n * comps
        assert data[pos:pos + 1] == b"\n"
        pos += 1
        arrays[kind[1]] = a.reshape(n, comps) if comps == 3 else a
    return hdr, arrays
