"""Python host of the MI355X solver: the reference's `SolverGPU` surface (src/solver/SolverGPU.ts)
and `LbmSolver` contract (src/backend/types.ts:7-18) over the C ABI.

Method names follow the reference (camelCase) so host code and tests read like the reference's.
Parameters are given as a dict shaped like `SolverParams` (src/solver/types.ts:3-29).
The reference's stepAndSample is a Promise; here it is a blocking call returning the stats.
"""
from dataclasses import dataclass

import numpy as np

from . import _native as N

D2Q9_C = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0],
                   [1, 1, 0], [-1, 1, 0], [-1, -1, 0], [1, -1, 0]], dtype=np.int64)
D2Q9_W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
D3Q19_C = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                    [1, 1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, -1], [0, 1, 1], [0, -1, -1],
                    [1, -1, 0], [-1, 1, 0], [1, 0, -1], [-1, 0, 1], [0, 1, -1], [0, -1, 1]], dtype=np.int64)
D3Q19_W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


@dataclass
class Lattice:
    """Velocity set (src/solver/lattice.ts:5-60)."""
    name: str
    dim: int
    q: int
    c: np.ndarray
    w: np.ndarray


D2Q9 = Lattice("D2Q9", 2, 9, D2Q9_C, D2Q9_W)
D3Q19 = Lattice("D3Q19", 3, 19, D3Q19_C, D3Q19_W)


@dataclass
class SolverStats:
    """SolverStats (src/solver/types.ts:31-42)."""
    fx: float
    fy: float
    fz: float
    maxU: float
    rhoRef: float
    unstable: bool
    step: int


@dataclass
class FieldExport:
    """FieldExport (src/solver/types.ts:45-55): arrays of shape (points, 4), x fastest."""
    dims: tuple
    stride: int
    vel: np.ndarray   # (ux, uy, uz, rho)
    mean: np.ndarray  # time-averaged (ux, uy, uz, rho)
    vort: np.ndarray  # (wx, wy, wz, Q)


DEFAULTS = dict(u=0.05, nu=0.01, cs=0.1, ground="freeslip", sides="freeslip", spongeNu=0.0, spongeStart=0.9,
                spongeIn=0.0, tauWall=0.5, precision="f32", interp=True, collision="regularized",
                closure="smagorinsky", emaAlpha=0.0)

_FIELDS = [  # (SolverParams key, struct field, mask bit, encoder)
    ("u", "u", 0, float), ("nu", "nu", 1, float), ("cs", "cs", 2, float),
    ("ground", "ground", 3, lambda v: N.GROUND[v]), ("sides", "sides", 4, lambda v: N.SIDES[v]),
    ("spongeNu", "sponge_nu", 5, float), ("spongeStart", "sponge_start", 6, float),
    ("spongeIn", "sponge_in", 7, float), ("tauWall", "tau_wall", 8, float),
    ("interp", "interp", 9, lambda v: 1 if v else 0),
    ("collision", "collision", 10, lambda v: N.COLLISION[v]), ("closure", "closure", 11, lambda v: N.CLOSURE[v]),
    ("emaAlpha", "ema_alpha", 12, float),
]


def _params_struct(p):
    s = N.Params()
    mask = 0
    for key, field, bit, enc in _FIELDS:
        if key in p and p[key] is not None:
            setattr(s, field, enc(p[key]))
            mask |= 1 << bit
    s.precision = N.PRECISION[p.get("precision") or "f32"]
    return s, mask


class SolverGPU:
    """GPU lattice-Boltzmann solver (D2Q9 when nz == 1, D3Q19 otherwise), SolverGPU.ts:12-444."""

    def __init__(self, dims, params):
        self._lib = N.load()
        if isinstance(dims, dict):
            nx, ny, nz = dims["nx"], dims["ny"], dims["nz"]
        else:
            nx, ny, nz = dims
        self.dims = dict(nx=int(nx), ny=int(ny), nz=int(nz))
        self.is3D = nz > 1
        self.lattice = D3Q19 if self.is3D else D2Q9
        self.n = nx * ny * nz
        p = dict(DEFAULTS)
        p.update({k: v for k, v in params.items() if v is not None})
        self._p = p
        ps, _ = _params_struct(p)
        d = N.Dims(nx, ny, nz)
        h = N.C.c_void_p()
        N.check(self._lib.wt_create(N.C.byref(d), N.C.byref(ps), N.C.byref(h)))
        self._h = h
        N.track(self)
        info = self.info()
        self.precision = N.PRECISION_NAME[info.precision]
        self.packed = self.precision != "f32"

    # -- lifecycle --------------------------------------------------------------------------
    def destroy(self):
        if getattr(self, "_h", None):
            self._lib.wt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    @staticmethod
    def bytesFor(dims):
        d = N.Dims(dims["nx"], dims["ny"], dims["nz"])
        return int(N.load().wt_bytes_for(N.C.byref(d)))

    def info(self):
        i = N.Info()
        N.check(self._lib.wt_info_get(self._h, N.C.byref(i)))
        return i

    @property
    def stepCount(self):
        return int(self.info().step)

    @property
    def params_(self):
        return dict(self._p)

    @property
    def tau0(self):
        return float(self._lib.wt_tau0(self._h))

    def setParams(self, p):
        self._p.update(p)
        ps, mask = _params_struct(p)
        N.check(self._lib.wt_set_params(self._h, N.C.byref(ps), mask))

    def reset(self):
        N.check(self._lib.wt_reset(self._h))

    def clearStats(self):
        N.check(self._lib.wt_clear_stats(self._h))

    # -- geometry ---------------------------------------------------------------------------
    def uploadFlags(self, flags):
        a = np.ascontiguousarray(flags, dtype=np.uint32).reshape(-1)
        if a.size != self.n:
            raise ValueError("flags must have n = %d entries, got %d" % (self.n, a.size))
        N.check(self._lib.wt_upload_flags(self._h, N.ptr(a)))

    def readFlags(self):
        out = np.empty(self.n, dtype=np.uint32)
        N.check(self._lib.wt_read_flags(self._h, N.ptr(out)))
        return out

    def uploadSdf(self, d):
        a = np.ascontiguousarray(d, dtype=np.float32).reshape(-1)
        if a.size != self.n:
            raise ValueError("sdf must have n = %d entries, got %d" % (self.n, a.size))
        N.check(self._lib.wt_upload_sdf(self._h, N.ptr(a)))

    def readSdf(self):
        """Stored wall distances (cells), -1 where unknown."""
        out = np.empty(self.n, dtype=np.float32)
        N.check(self._lib.wt_read_sdf(self._h, N.ptr(out)))
        return out

    def setWheels(self, wheels):
        """wheels: list of dicts {c: [x,y,z], axis: [x,y,z], r, hw} in grid space (SolverGPU.ts:325)."""
        a = wheel_array(wheels)
        N.check(self._lib.wt_set_wheels(self._h, N.ptr(a) if len(a) else None, len(a) // 8))
        self._wheels = [dict(w) for w in list(wheels or [])[:16]]

    @property
    def wheels_(self):
        """The wheel list last set (setWheels or voxelizeMesh)."""
        return list(getattr(self, "_wheels", []))

    def voxelizeSdf(self, prims):
        """Analytic-SDF voxelizer (a list of N.Prim); returns VoxelInfo as a dict."""
        arr = (N.Prim * max(len(prims), 1))(*prims)
        vi = N.VoxelInfo()
        N.check(self._lib.wt_voxelize_sdf(self._h, arr, len(prims), N.C.byref(vi)))
        return voxel_info(vi)

    def voxelizeMesh(self, positions, indices, matrix, wheels=()):
        """VoxelizerGPU.voxelize (src/voxelize/VoxelizerGPU.ts:48-151). matrix: 16 floats, column-major."""
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
        idx = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
        m = np.ascontiguousarray(matrix, dtype=np.float32).reshape(16)
        wa = wheel_array(wheels)
        vi = N.VoxelInfo()
        N.check(self._lib.wt_voxelize_mesh(self._h, N.ptr(pos), pos.size // 3, N.ptr(idx), idx.size // 3, N.ptr(m),
                                           N.ptr(wa) if len(wa) else None, len(wa) // 8, N.C.byref(vi)))
        self._wheels = [dict(w) for w in list(wheels or [])[:16]]
        return voxel_info(vi)

    # -- stepping ---------------------------------------------------------------------------
    def encodeSteps(self, steps):
        """Enqueue `steps` steps (asynchronous), SolverGPU.ts:204-222."""
        N.check(self._lib.wt_step(self._h, int(steps)))

    step = encodeSteps

    def stepAndSample(self, steps):
        """Run `steps` steps, then macro + force reduction, and return SolverStats (SolverGPU.ts:292-309)."""
        st = N.Stats()
        N.check(self._lib.wt_step_and_sample(self._h, int(steps), N.C.byref(st)))
        return SolverStats(st.fx, st.fy, st.fz, st.max_u, st.rho_ref, bool(st.unstable), int(st.step))

    def encodePost(self, derived=True):
        """Macro pass, force reduction and (derived) the derived pass, with an asynchronous read of the
        statistics queued (SolverGPU.ts:225-261). Returns True when a read was queued."""
        q = N.C.c_int32()
        N.check(self._lib.wt_encode_post(self._h, 1 if derived else 0, N.C.byref(q)))
        return bool(q.value)

    def flushReads(self, onStats=None):
        """Non-blocking: the statistics of the queued reads that have completed, in posting order
        (SolverGPU.ts:264-287); each goes to onStats and the latest by step is kept in self.latest.
        Returns (list of SolverStats, reads still pending)."""
        arr = (N.Stats * 4)()
        cnt, pend = N.C.c_int32(), N.C.c_int32()
        N.check(self._lib.wt_flush_reads(self._h, arr, 4, N.C.byref(cnt), N.C.byref(pend)))
        out = [SolverStats(a.fx, a.fy, a.fz, a.max_u, a.rho_ref, bool(a.unstable), int(a.step)) for a in arr[:cnt.value]]
        for s in out:
            if getattr(self, "latest", None) is None or s.step >= self.latest.step:
                self.latest = s
            if onStats:
                onStats(s)
        return out, pend.value

    def sampleRaw(self):
        out = np.zeros(8, dtype=np.float64)
        N.check(self._lib.wt_sample_raw(self._h, N.ptr(out)))
        return out

    def sync(self):
        N.check(self._lib.wt_sync(self._h))

    def fence(self):
        """A ticket for the point after all work queued on this solver so far (the queue's
        onSubmittedWorkDone, backend/webgpu.ts:120, which the reference's frame throttle counts)."""
        t = N.C.c_int64()
        N.check(self._lib.wt_fence(self._h, N.C.byref(t)))
        return t.value

    def fenceDone(self, ticket):
        """Non-blocking: True once the GPU has passed the fence."""
        d = N.C.c_int32()
        N.check(self._lib.wt_fence_done(self._h, int(ticket), N.C.byref(d)))
        return bool(d.value)

    def fenceWait(self, ticket):
        N.check(self._lib.wt_fence_wait(self._h, int(ticket)))

    # -- readback ---------------------------------------------------------------------------
    def readPopulations(self):
        """Current populations, unpacked to full f_i, shape (q, n) float32."""
        out = np.empty((self.lattice.q, self.n), dtype=np.float32)
        N.check(self._lib.wt_read_populations(self._h, N.ptr(out)))
        return out

    def readPopulationPlanes(self, z0, count):
        """Populations of the planes z0 .. z0 + count - 1, unpacked, shape (q, nx * ny * count) float32
        (x fastest, then y, then z - z0)."""
        m = self.dims["nx"] * self.dims["ny"] * int(count)
        out = np.empty((self.lattice.q, m), dtype=np.float32)
        N.check(self._lib.wt_read_population_planes(self._h, int(z0), int(count), N.ptr(out)))
        return out

    def uploadPopulations(self, f):
        a = np.ascontiguousarray(f, dtype=np.float32).reshape(self.lattice.q, self.n)
        N.check(self._lib.wt_upload_populations(self._h, N.ptr(a)))

    def readMacro(self):
        vel = np.empty((self.n, 4), dtype=np.float32)
        ma = np.empty((self.n, 4), dtype=np.float32)
        mb = np.empty(self.n, dtype=np.float32)
        N.check(self._lib.wt_read_macro(self._h, N.ptr(vel), N.ptr(ma), N.ptr(mb)))
        return vel, ma, mb

    # -- derived fields and export ------------------------------------------------------------
    def derive(self):
        """Derived pass of encodePost(derived=True) (SolverGPU.ts:242-246, wgslLbm.ts:527-592) over
        the fields of the last sample: vorticity + Q-criterion, turbulence intensity, mean copy."""
        N.check(self._lib.wt_derive(self._h))

    def readDerived(self):
        """(vort (n, 4) = wx, wy, wz, Q; stat (n, 4) = TI, solid, 0, 0; mean (n, 4)) as float32."""
        out = [np.empty((self.n, 4), dtype=np.float32) for _ in range(3)]
        N.check(self._lib.wt_read_derived(self._h, *(N.ptr(a) for a in out)))
        return tuple(out)

    def exportFields(self, stride):
        """SolverGPU.exportFields (SolverGPU.ts:374-425): the velocity, mean and vorticity fields
        sub-sampled by `stride` in every direction."""
        dims = (N.C.c_int32 * 3)()
        N.check(self._lib.wt_export_dims(self._h, int(stride), dims))
        m = int(dims[0]) * int(dims[1]) * int(dims[2])
        vel, mean, vort = (np.empty((m, 4), dtype=np.float32) for _ in range(3))
        N.check(self._lib.wt_export_fields(self._h, int(stride), N.ptr(vel), N.ptr(mean), N.ptr(vort), dims))
        return FieldExport((int(dims[0]), int(dims[1]), int(dims[2])), int(stride), vel, mean, vort)

    # -- probes, streamlines, tracers (render/tracers.ts) --------------------------------------
    def probe(self, points):
        """Field and mean field at points (k, 3) in cell units: returns (vel (k, 4), mean (k, 4))."""
        p = np.zeros((len(points), 4), np.float32)
        p[:, :3] = np.asarray(points, dtype=np.float32).reshape(-1, 3)
        out = np.empty((len(p), 8), np.float32)
        N.check(self._lib.wt_probe(self._h, N.ptr(p), len(p), N.ptr(out)))
        return out[:, :4], out[:, 4:]

    def streamlines(self, rake, seeds, points, h=0.5, perLine=None, useMean=False):
        """(seeds, points, 4) array of (x, y, z, speed or -1) along streamlines seeded on `rake`
        (dict a, b, d of 3 floats; perLine seeds per rake line, default all on one line)."""
        r = rake_struct(rake)
        out = np.empty((seeds, points, 4), np.float32)
        N.check(self._lib.wt_streamlines(self._h, N.C.byref(r), int(seeds), int(perLine or seeds), int(points), float(h),
                                         1 if useMean else 0, N.ptr(out)))
        return out

    # -- profiling --------------------------------------------------------------------------
    PROFILE_MODES = {False: 0, True: 1, "off": 0, "launch": 1, "span": 2}

    def profileEnable(self, on=True):
        """Step-kernel timing: True / "launch" binds HIP events to every launch (exact durations,
        a few microseconds of queue time added between launches); "span" times each step() call's
        launches as one stretch of the stream (no events between launches; the per-launch average
        includes the gaps between back-to-back launches); False turns it off."""
        mode = self.PROFILE_MODES.get(on) if isinstance(on, (bool, str)) else int(on)
        if mode is None:
            raise ValueError("profile mode must be one of %s" % sorted(k for k in self.PROFILE_MODES if isinstance(k, str)))
        N.check(self._lib.wt_profile_enable(self._h, mode))

    def profileRead(self):
        b = N.C.c_double()
        k = N.C.c_int64()
        bd = N.C.c_double()
        N.check(self._lib.wt_profile_read(self._h, N.C.byref(b), N.C.byref(k), N.C.byref(bd)))
        return b.value, k.value, bd.value

    def setTuning(self, key, value):
        N.check(self._lib.wt_set_tuning(self._h, int(key), int(value)))

    def getTuning(self, key):
        v = N.C.c_int32()
        N.check(self._lib.wt_get_tuning(self._h, int(key), N.C.byref(v)))
        return v.value

    def autotune(self, steps=20):
        """Time every candidate launch configuration for `steps` steps, keep the fastest (the state
        is restored afterwards); returns {chunk, cellsPerThread, msPerStep, trials: [...]}."""
        r = N.TuneResult()
        N.check(self._lib.wt_autotune(self._h, int(steps), N.C.byref(r)))
        trials = [dict(chunk=r.trial_chunk[k], cellsPerThread=r.trial_cpt[k], msPerStep=r.trial_ms[k])
                  for k in range(r.trials)]
        return dict(chunk=r.chunk, cellsPerThread=r.cells_per_thread, msPerStep=r.ms_per_step, trials=trials)

    def workSplit(self):
        a = N.C.c_int64()
        b = N.C.c_int64()
        N.check(self._lib.wt_work_split(self._h, N.C.byref(a), N.C.byref(b)))
        return a.value, b.value


def rake_struct(rake):
    r = N.Rake()
    for k in ("a", "b", "d"):
        v = list(rake.get(k, (0.0, 0.0, 0.0)))[:3] + [0.0]
        getattr(r, k)[:] = [float(x) for x in v]
    return r


class Tracers:
    """Particle tracers on a solver's velocity field (PARTICLES_WGSL, render/tracers.ts:3-106)."""

    def __init__(self, solver, count, trail=16, emitter="inlet", nozzles=16, lines=1, maxAge=400.0, rake=None):
        self.solver = solver
        self.count, self.trail = int(count), int(trail)
        o = N.ParticleOpts()
        o.count, o.trail = self.count, self.trail
        o.emitter = 0 if emitter == "inlet" else 1
        o.nozzles, o.lines, o.max_age = int(nozzles), int(lines), float(maxAge)
        o.rake = rake_struct(rake or {})
        self._lib = solver._lib
        h = N.C.c_void_p()
        N.check(self._lib.wt_particles_create(solver._h, N.C.byref(o), N.C.byref(h)))
        self._h = h
        N.track(self)
        self.frameSeed = 0

    def update(self, steps, advance=True, fill=False, seed=None):
        """One frame: `steps` lattice steps of advection (the renderer passes a frame counter as seed)."""
        if seed is None:
            seed = self.frameSeed
            self.frameSeed += 1
        N.check(self._lib.wt_particles_update(self._h, float(steps), int(seed) & 0xFFFFFFFF, 1 if fill else 0, 1 if advance else 0))

    def read(self):
        """(state (count, 4) = x, y, z, age; trail (trail, count, 4); head slot)."""
        st = np.empty((self.count, 4), np.float32)
        hist = np.empty((self.trail, self.count, 4), np.float32)
        head = N.C.c_int32()
        N.check(self._lib.wt_particles_read(self._h, N.ptr(st), N.ptr(hist), N.C.byref(head)))
        return st, hist, head.value

    def destroy(self):
        if getattr(self, "_h", None):
            N.check(self._lib.wt_particles_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def wheel_array(wheels):
    out = []
    for w in list(wheels)[:16] if wheels is not None else []:
        out += [w["c"][0], w["c"][1], w["c"][2], w["axis"][0], w["axis"][1], w["axis"][2], w["r"], w["hw"]]
    return np.array(out, dtype=np.float32)


def voxel_info(vi):
    return dict(frontal=int(vi.frontal), solidCells=int(vi.solid_cells), min=list(vi.min), max=list(vi.max))


class BatchGPU:
    """Many independent SolverGPU domains stepped by one launch per step (wt_batch_*), in
    hipGraph-captured chunks. The solvers stay owned by the caller; destroy the batch first."""

    def __init__(self, solvers, chunk=100, graph=True):
        self._lib = N.load()
        self.solvers = list(solvers)
        arr = (N.C.c_void_p * len(self.solvers))(*[s._h.value for s in self.solvers])
        h = N.C.c_void_p()
        N.check(self._lib.wt_batch_create(arr, len(self.solvers), N.C.byref(h)))
        self._h = h
        N.track(self)
        self.setChunk(chunk, graph)

    def setChunk(self, chunk, graph=True):
        N.check(self._lib.wt_batch_set_chunk(self._h, int(chunk), 1 if graph else 0))

    def run(self, steps):
        N.check(self._lib.wt_batch_run(self._h, int(steps)))

    def sync(self):
        N.check(self._lib.wt_batch_sync(self._h))

    def stats(self):
        """SolverStats per member (synchronises)."""
        arr = (N.Stats * len(self.solvers))()
        N.check(self._lib.wt_batch_stats(self._h, arr))
        return [SolverStats(a.fx, a.fy, a.fz, a.max_u, a.rho_ref, bool(a.unstable), int(a.step)) for a in arr]

    def destroy(self):
        if getattr(self, "_h", None):
            self._lib.wt_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
