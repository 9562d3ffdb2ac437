"""ctypes binding of libwt.so (the C ABI declared in include/wt.h).

The library is loaded from the package's in-tree build (windtunnel_vis_amd/_lib/libwt.so). There
is no fallback: if the library is missing or cannot be loaded, `load()` raises.
"""
import atexit
import ctypes as C
import os
import weakref

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libwt.so")

WT_OK = 0
WT_E_ARG = -1
WT_E_HIP = -2
WT_E_NOMEM = -3
WT_E_STATE = -4
WT_E_NODEV = -5

GROUND = {"freeslip": 0, "noslip": 1, "moving": 2}
SIDES = {"freeslip": 0, "periodic": 1}
PRECISION = {"f32": 0, "f16": 1, "bf16": 2}
PRECISION_NAME = {v: k for k, v in PRECISION.items()}
COLLISION = {"regularized": 0, "bgk": 1}
CLOSURE = {"smagorinsky": 0, "mlp": 1}

TUNE_ITEM_CHUNK = 0
TUNE_CELLS_PER_THREAD = 1
TUNE_MAX_TRIALS = 16

PRIM_ROUNDED_BOX = 1
PRIM_CYLINDER = 2
PRIM_SPHERE = 3
PRIM_WEDGE = 4


class Dims(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32)]


class Params(C.Structure):
    _fields_ = [
        ("u", C.c_double), ("nu", C.c_double), ("cs", C.c_double),
        ("ground", C.c_int32), ("sides", C.c_int32),
        ("sponge_nu", C.c_double), ("sponge_start", C.c_double), ("sponge_in", C.c_double),
        ("tau_wall", C.c_double),
        ("precision", C.c_int32), ("interp", C.c_int32), ("collision", C.c_int32), ("closure", C.c_int32),
        ("ema_alpha", C.c_double),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("fx", C.c_double), ("fy", C.c_double), ("fz", C.c_double), ("max_u", C.c_double),
        ("rho_ref", C.c_double), ("unstable", C.c_int32), ("pad", C.c_int32), ("step", C.c_int64),
    ]


class Info(C.Structure):
    _fields_ = [
        ("dims", Dims), ("n", C.c_int64), ("q", C.c_int32), ("precision", C.c_int32), ("step", C.c_int64),
        ("wall_cells", C.c_int64), ("device_bytes", C.c_int64),
    ]


class VoxelInfo(C.Structure):
    _fields_ = [("frontal", C.c_int64), ("solid_cells", C.c_int64), ("min", C.c_int32 * 3), ("max", C.c_int32 * 3)]


class TuneResult(C.Structure):
    _fields_ = [("chunk", C.c_int32), ("cells_per_thread", C.c_int32), ("ms_per_step", C.c_double),
                ("trials", C.c_int32), ("pad", C.c_int32), ("trial_chunk", C.c_int32 * 16),
                ("trial_cpt", C.c_int32 * 16), ("trial_ms", C.c_double * 16)]


MLP_NPARAM = 1377


class Mlp(C.Structure):
    _fields_ = [("params", C.c_float * MLP_NPARAM), ("mu", C.c_float * 8), ("sigma", C.c_float * 8), ("nu_scale", C.c_float),
                ("smag_floor", C.c_float)]


class TrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_double), ("beta1", C.c_double),
                ("beta2", C.c_double), ("eps", C.c_double), ("seed", C.c_uint64), ("max_samples", C.c_int64)]


class Prim(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("flags", C.c_int32), ("center", C.c_float * 3), ("rot", C.c_float * 9),
        ("half", C.c_float * 3), ("radius", C.c_float), ("param", C.c_float),
    ]


# every exported symbol of include/wt.h with its signature: (restype, argtypes)
_P = C.c_void_p
class Rake(C.Structure):
    _fields_ = [("a", C.c_float * 4), ("b", C.c_float * 4), ("d", C.c_float * 4)]


class ParticleOpts(C.Structure):
    _fields_ = [("count", C.c_int32), ("trail", C.c_int32), ("emitter", C.c_int32), ("nozzles", C.c_int32),
                ("lines", C.c_int32), ("max_age", C.c_float), ("rake", Rake)]


SIGNATURES = {
    "wt_create": (C.c_int, [C.POINTER(Dims), C.POINTER(Params), C.POINTER(_P)]),
    "wt_destroy": (C.c_int, [_P]),
    "wt_bytes_for": (C.c_int64, [C.POINTER(Dims)]),
    "wt_info_get": (C.c_int, [_P, C.POINTER(Info)]),
    "wt_last_error": (C.c_char_p, []),
    "wt_set_params": (C.c_int, [_P, C.POINTER(Params), C.c_uint32]),
    "wt_get_params": (C.c_int, [_P, C.POINTER(Params)]),
    "wt_tau0": (C.c_double, [_P]),
    "wt_reset": (C.c_int, [_P]),
    "wt_clear_stats": (C.c_int, [_P]),
    "wt_upload_flags": (C.c_int, [_P, _P]),
    "wt_read_flags": (C.c_int, [_P, _P]),
    "wt_upload_sdf": (C.c_int, [_P, _P]),
    "wt_read_sdf": (C.c_int, [_P, _P]),
    "wt_set_wheels": (C.c_int, [_P, _P, C.c_int32]),
    "wt_voxelize_sdf": (C.c_int, [_P, C.POINTER(Prim), C.c_int32, C.POINTER(VoxelInfo)]),
    "wt_voxelize_mesh": (C.c_int, [_P, _P, C.c_int32, _P, C.c_int32, _P, _P, C.c_int32, C.POINTER(VoxelInfo)]),
    "wt_step": (C.c_int, [_P, C.c_int32]),
    "wt_step_and_sample": (C.c_int, [_P, C.c_int32, C.POINTER(Stats)]),
    "wt_sample_raw": (C.c_int, [_P, _P]),
    "wt_step_and_sample_async": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int64)]),
    "wt_sample_wait": (C.c_int, [_P, C.c_int64, C.POINTER(Stats)]),
    "wt_encode_post": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32)]),
    "wt_flush_reads": (C.c_int, [_P, C.POINTER(Stats), C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "wt_sync": (C.c_int, [_P]),
    "wt_fence": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "wt_fence_done": (C.c_int, [_P, C.c_int64, C.POINTER(C.c_int32)]),
    "wt_fence_wait": (C.c_int, [_P, C.c_int64]),
    "wt_read_populations": (C.c_int, [_P, _P]),
    "wt_upload_populations": (C.c_int, [_P, _P]),
    "wt_read_population_planes": (C.c_int, [_P, C.c_int32, C.c_int32, _P]),
    "wt_read_macro": (C.c_int, [_P, _P, _P, _P]),
    "wt_derive": (C.c_int, [_P]),
    "wt_read_derived": (C.c_int, [_P, _P, _P, _P]),
    "wt_export_dims": (C.c_int, [_P, C.c_int32, _P]),
    "wt_export_fields": (C.c_int, [_P, C.c_int32, _P, _P, _P, _P]),
    "wt_probe": (C.c_int, [_P, _P, C.c_int32, _P]),
    "wt_streamlines": (C.c_int, [_P, C.POINTER(Rake), C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_int32, _P]),
    "wt_particles_create": (C.c_int, [_P, C.POINTER(ParticleOpts), C.POINTER(_P)]),
    "wt_particles_destroy": (C.c_int, [_P]),
    "wt_particles_update": (C.c_int, [_P, C.c_float, C.c_uint32, C.c_int32, C.c_int32]),
    "wt_particles_read": (C.c_int, [_P, _P, _P, C.POINTER(C.c_int32)]),
    "wt_profile_enable": (C.c_int, [_P, C.c_int32]),
    "wt_profile_read": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "wt_work_split": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "wt_copy_probe": (C.c_int, [C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
    "wt_set_tuning": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "wt_get_tuning": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_int32)]),
    "wt_autotune": (C.c_int, [_P, C.c_int32, C.POINTER(TuneResult)]),
    "wt_batch_create": (C.c_int, [C.POINTER(_P), C.c_int32, C.POINTER(_P)]),
    "wt_batch_destroy": (C.c_int, [_P]),
    "wt_batch_set_chunk": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "wt_batch_run": (C.c_int, [_P, C.c_int32]),
    "wt_batch_sync": (C.c_int, [_P]),
    "wt_batch_stats": (C.c_int, [_P, C.POINTER(Stats)]),
    "wt_mlp_load": (C.c_int, [_P, C.POINTER(Mlp)]),
    "wt_read_wall_distance": (C.c_int, [_P, C.c_void_p]),
    "wt_mlp_infer": (C.c_int, [C.POINTER(Mlp), C.c_void_p, C.c_int64, C.c_void_p]),
    "wt_mlp_infer_noalloc": (C.c_int, [C.POINTER(Mlp), C.c_void_p, C.c_int64, C.c_void_p]),
    "wt_dataset_create": (C.c_int, [C.c_int64, C.POINTER(_P)]),
    "wt_dataset_destroy": (C.c_int, [_P]),
    "wt_dataset_add": (C.c_int, [_P, _P, _P, C.POINTER(C.c_int64)]),
    "wt_dataset_count": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "wt_dataset_read": (C.c_int, [_P, C.c_int64, C.c_int64, C.c_void_p]),
    "wt_dataset_append": (C.c_int, [_P, C.c_void_p, C.c_int64]),
    "wt_mlp_train": (C.c_int, [_P, C.POINTER(Mlp), C.POINTER(TrainOpts), C.c_void_p]),
    "wt_mlp_grad": (C.c_int, [_P, C.POINTER(Mlp), C.c_int64, C.c_int64, C.c_void_p, C.POINTER(C.c_double)]),
}

_lib = None


class WtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("libwt error %d: %s" % (code, msg))
        self.code = code


def load():
    """Load libwt.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libwt.so is not built (%s); run `python -c 'import __graft_entry__ as g; g.build()'`" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code):
    if code != WT_OK:
        msg = _lib.wt_last_error().decode() if _lib is not None else ""
        raise WtError(code, msg)
    return code


def ptr(a):
    """Raw data pointer of a C-contiguous numpy array."""
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return C.c_void_p(a.ctypes.data)


# Handles still open when the interpreter exits are destroyed by an atexit hook, newest first (a
# tracer before its solver), while the HIP runtime is certainly alive: left to the garbage
# collector, they could be finalized after the runtime's own teardown.
_live = []


def track(obj):
    if len(_live) > 256:
        _live[:] = [r for r in _live if r() is not None]
    _live.append(weakref.ref(obj))


@atexit.register
def _close_all():
    while _live:
        obj = _live.pop()()
        if obj is not None:
            try:
                obj.destroy()
            except Exception:
                pass

