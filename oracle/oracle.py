"""Python wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/_build/liboracle.so (built from oracle/lbm_oracle.cpp by oracle/Makefile) and exposes
the same surface as the product's SolverGPU so parity tests can drive both with one script.
Only tests/, the smoke check and bench.py's CPU-baseline leg may import this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

GROUND = {"freeslip": 0, "noslip": 1, "moving": 2}
SIDES = {"freeslip": 0, "periodic": 1}
STORAGE = {None: 0, "f64": 0, "f32": 1, "f16": 2, "bf16": 3}
COLLISION = {"regularized": 0, "bgk": 1}


class OrParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("u", C.c_double), ("nu", C.c_double), ("cs", C.c_double),
        ("ground", C.c_int32), ("sides", C.c_int32),
        ("sponge_nu", C.c_double), ("sponge_start", C.c_double), ("sponge_in", C.c_double), ("tau_wall", C.c_double),
        ("interp", C.c_int32), ("collision", C.c_int32), ("ema_alpha", C.c_double),
        ("storage", C.c_int32), ("periodic_x", C.c_int32), ("threads", C.c_int32),
    ]


class OrMlp(C.Structure):
    _fields_ = [("params", C.c_double * 1377), ("mu", C.c_double * 8), ("sigma", C.c_double * 8),
                ("nu_scale", C.c_double), ("smag_floor", C.c_double), ("pair", C.c_int32), ("pad", C.c_int32)]


def mlp_struct(net, pair=2):
    """OrMlp from a closure network (windtunnel_vis_amd.closure.Network or anything with params,
    mu, sigma, nu_scale, smag_floor)."""
    m = OrMlp()
    m.params[:] = [float(v) for v in np.asarray(net.params, np.float64)]
    m.mu[:] = [float(v) for v in np.asarray(net.mu, np.float64)]
    m.sigma[:] = [float(v) for v in np.asarray(net.sigma, np.float64)]
    m.nu_scale = float(net.nu_scale)
    m.smag_floor = float(net.smag_floor)
    m.pair = int(pair)
    return m


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sig = {
        "or_create": (P, [C.POINTER(OrParams), C.c_int]),
        "or_destroy": (None, [P]),
        "or_q": (C.c_int, [P]),
        "or_set_params": (None, [P, C.POINTER(OrParams)]),
        "or_upload_flags": (None, [P, P]),
        "or_upload_sdf": (None, [P, P]),
        "or_set_wheels": (None, [P, P, C.c_int]),
        "or_reset": (None, [P]),
        "or_clear_stats": (None, [P]),
        "or_step": (None, [P, C.c_int]),
        "or_sample": (None, [P, P]),
        "or_step_count": (C.c_int64, [P]),
        "or_read_populations": (None, [P, P]),
        "or_upload_populations": (None, [P, P]),
        "or_read_macro": (None, [P, P, P, P]),
        "or_round_half": (C.c_double, [C.c_double]),
        "or_round_bf16": (C.c_double, [C.c_double]),
        "or_set_mlp": (None, [P, P, P]),
        "or_mlp_infer": (None, [P, P, C.c_int64, P]),
        "or_mlp_margins": (None, [P, P]),
        "or_mlp_features": (None, [P, P, P, C.c_int64, C.c_double, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def _params(dims, p, storage, periodic_x, threads):
    s = OrParams()
    s.nx, s.ny, s.nz = dims
    s.u = p.get("u", 0.05)
    s.nu = p.get("nu", 0.01)
    s.cs = p.get("cs", 0.1)
    s.ground = GROUND[p.get("ground", "freeslip")]
    s.sides = SIDES[p.get("sides", "freeslip")]
    s.sponge_nu = p.get("spongeNu", 0.0)
    s.sponge_start = p.get("spongeStart", 0.9)
    s.sponge_in = p.get("spongeIn") or 0.0
    s.tau_wall = p.get("tauWall") if p.get("tauWall") is not None else 0.5
    s.interp = 0 if p.get("interp") is False else 1
    s.collision = COLLISION[p.get("collision") or "regularized"]
    s.ema_alpha = p.get("emaAlpha", 0.0)
    s.storage = STORAGE[storage]
    s.periodic_x = 1 if periodic_x else 0
    s.threads = threads
    return s


class OracleSolver:
    """FP64 (or FP32) CPU restatement with the SolverGPU surface.

    storage: None/'f64' (no rounding), 'f32', 'f16' or 'bf16' (deviation rounding at every store).
    """

    def __init__(self, dims, params, storage=None, periodic_x=False, fp32=False, threads=0):
        self._lib = load()
        if isinstance(dims, dict):
            dims = (dims["nx"], dims["ny"], dims["nz"])
        self.dims = dict(nx=dims[0], ny=dims[1], nz=dims[2])
        self._dims = tuple(dims)
        self.n = dims[0] * dims[1] * dims[2]
        self._p = dict(params)
        self._storage = storage
        self._px = periodic_x
        self._threads = threads
        ps = _params(self._dims, self._p, storage, periodic_x, threads)
        self._h = self._lib.or_create(C.byref(ps), 1 if fp32 else 0)
        self.q = self._lib.or_q(self._h)

    def destroy(self):
        if self._h:
            self._lib.or_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    @property
    def stepCount(self):
        return int(self._lib.or_step_count(self._h))

    def setParams(self, p):
        self._p.update(p)
        ps = _params(self._dims, self._p, self._storage, self._px, self._threads)
        self._lib.or_set_params(self._h, C.byref(ps))

    def reset(self):
        self._lib.or_reset(self._h)

    def clearStats(self):
        self._lib.or_clear_stats(self._h)

    def uploadFlags(self, flags):
        a = np.ascontiguousarray(flags, dtype=np.uint32).reshape(-1)
        self._lib.or_upload_flags(self._h, _ptr(a))

    def uploadSdf(self, d):
        a = np.ascontiguousarray(d, dtype=np.float32).reshape(-1)
        self._lib.or_upload_sdf(self._h, _ptr(a))

    def setWheels(self, wheels):
        out = []
        for w in list(wheels)[:16]:
            out += [w["c"][0], w["c"][1], w["c"][2], w["axis"][0], w["axis"][1], w["axis"][2], w["r"], w["hw"]]
        a = np.array(out, dtype=np.float32)
        self._lib.or_set_wheels(self._h, _ptr(a) if len(a) else None, len(a) // 8)

    def step(self, steps):
        self._lib.or_step(self._h, int(steps))

    encodeSteps = step

    def sample(self):
        """Run the macro pass + reduction (updates the EMA statistics) and return the raw result."""
        out = np.zeros(8, dtype=np.float64)
        self._lib.or_sample(self._h, _ptr(out))
        self._last = out
        return out

    def sampleRaw(self):
        """The raw result of the last sample (like wt_sample_raw)."""
        return getattr(self, "_last", np.zeros(8))

    def stepAndSample(self, steps):
        """Same decode as SolverGPU.ts:302-308."""
        from windtunnel_vis_amd.solver import SolverStats
        self.step(steps)
        d = self.sample()
        return SolverStats(d[0], d[1], d[2], float(np.sqrt(max(d[3], 0.0))), d[4] / d[5] if d[5] > 0 else 1.0,
                           bool(d[6] > 0 or not np.isfinite(d[0]) or d[3] > 0.25), self.stepCount)

    def readPopulations(self):
        out = np.empty((self.q, self.n), dtype=np.float64)
        self._lib.or_read_populations(self._h, _ptr(out))
        return out

    def uploadPopulations(self, f):
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(self.q, self.n)
        self._lib.or_upload_populations(self._h, _ptr(a))

    def readMacro(self):
        vel = np.empty((self.n, 4))
        ma = np.empty((self.n, 4))
        mb = np.empty(self.n)
        self._lib.or_read_macro(self._h, _ptr(vel), _ptr(ma), _ptr(mb))
        return vel, ma, mb

    def voxelizeMesh(self, positions, indices, matrix, wheels=()):
        """The mesh voxelizer's CPU restatement (voxelize_mesh) uploaded into this solver: flags
        with wheel tags, wall distances and the wheel list; returns the VoxelInfo dict."""
        flags, sdf_bits, info = voxelize_mesh(positions, indices, matrix, self._dims, wheels)
        self.setWheels(list(wheels))
        self.uploadFlags(flags)
        self.uploadSdf(decode_sdf(sdf_bits).astype(np.float32))
        return info

    def setMlp(self, net, wdist, pair=2):
        """Attach the learned closure: `net` a closure network, `wdist` the wall distance in cells
        (n values, stored as round(32 d) bytes like the product), `pair` the product's cells per
        lane (the closure runs on groups of `pair` cells along x with no wall or solid cell).
        net None detaches it."""
        if net is None:
            self._lib.or_set_mlp(self._h, None, None)
            return
        self._mlp = mlp_struct(net, pair)
        self._wd = np.ascontiguousarray(np.rint(np.asarray(wdist, np.float64).reshape(-1) * 32), dtype=np.uint8)
        self._lib.or_set_mlp(self._h, C.byref(self._mlp), _ptr(self._wd))


    def mlpMargins(self):
        """Per cell, the BF16 rounding margin of the last step's closure evaluation: the smallest
        distance of a BF16-rounded network input or positive first-layer value to a rounding
        midpoint (inf where the network did not run)."""
        out = np.empty(self.n)
        self._lib.or_mlp_margins(self._h, _ptr(out))
        return out


def mlp_infer(net, raw):
    """The closure network alone on raw inputs (n, 8): eddy viscosities (n,) in FP64 with the MFMA
    path's BF16 roundings."""
    r = np.ascontiguousarray(raw, dtype=np.float64).reshape(-1, 8)
    out = np.empty(len(r))
    m = mlp_struct(net)
    load().or_mlp_infer(C.byref(m), _ptr(r), len(r), _ptr(out))
    return out


def mlp_features(S6, wd, umag, nu0):
    """The eight raw closure inputs of n cells from their strain rates (n, 6), wall distances and speeds."""
    S = np.ascontiguousarray(S6, dtype=np.float64).reshape(-1, 6)
    n = len(S)
    w = np.ascontiguousarray(np.broadcast_to(np.asarray(wd, np.float64), (n,)))
    u = np.ascontiguousarray(np.broadcast_to(np.asarray(umag, np.float64), (n,)))
    out = np.empty((n, 8))
    load().or_mlp_features(_ptr(S), _ptr(w), _ptr(u), n, float(nu0), _ptr(out))
    return out


def round_half(x):
    return load().or_round_half(float(x))


def round_bf16(x):
    return load().or_round_bf16(float(x))


def voxelize_mesh(positions, indices, matrix, dims, wheels=()):
    """CPU restatement of the mesh voxelizer. Returns (flags u32[n], sdf_bits u32[n], info dict)."""
    lib = load()
    fn = lib.or_voxelize_mesh
    fn.restype = None
    P = C.c_void_p
    fn.argtypes = [P, C.c_int, P, C.c_int, P, P, C.c_int, C.c_int, C.c_int, C.c_int, P, P, P]
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
    idx = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
    m = np.ascontiguousarray(matrix, dtype=np.float32).reshape(16)
    wl = []
    for w in list(wheels)[:16]:
        wl += [w["c"][0], w["c"][1], w["c"][2], w["axis"][0], w["axis"][1], w["axis"][2], w["r"], w["hw"]]
    wa = np.array(wl if wl else [0.0], dtype=np.float32)
    nx, ny, nz = dims
    n = nx * ny * nz
    flags = np.zeros(n, np.uint32)
    sdf = np.zeros(n, np.uint32)
    info = np.zeros(8, np.int64)
    fn(_ptr(pos), pos.size // 3, _ptr(idx), idx.size // 3, _ptr(m), _ptr(wa), len(wl) // 8, nx, ny, nz, _ptr(flags), _ptr(sdf),
       _ptr(info))
    ok = info[1] > 0
    return flags, sdf, dict(frontal=int(info[0]), solidCells=int(info[1]),
                            min=[int(v) for v in info[2:5]] if ok else [0, 0, 0],
                            max=[int(v) + 1 for v in info[5:8]] if ok else [0, 0, 0])


def decode_sdf(bits):
    """~bits encoding -> distances (-1 unknown), SolverGPU.ts:312-322 / wgslLbm.ts:32-36."""
    b = np.asarray(bits, dtype=np.uint32)
    d = (~b).view(np.float32).astype(np.float64)
    return np.where(b == 0, -1.0, d)
