"""The C-ABI library builds, loads and exports every function include/wt.h declares (no GPU needed)."""
import ctypes
import os
import re

from windtunnel_vis_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "wt.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(wt_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_surface():
    names = declared_functions()
    for required in ["wt_create", "wt_destroy", "wt_step", "wt_step_and_sample", "wt_read_populations",
                     "wt_upload_flags", "wt_upload_sdf", "wt_set_wheels", "wt_voxelize_mesh", "wt_voxelize_sdf"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = N.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_signatures_cover_header():
    assert set(declared_functions()) == set(N.SIGNATURES)


def test_struct_layouts():
    # sizes the C side expects (natural alignment, x86-64)
    assert ctypes.sizeof(N.Dims) == 12
    assert ctypes.sizeof(N.Stats) == 56
    assert ctypes.sizeof(N.Prim) == 4 * 2 + 4 * 3 + 4 * 9 + 4 * 3 + 4 + 4


def test_errors_without_gpu_are_reported_not_crashes():
    lib = N.load()
    d = N.Dims(0, 0, 0)
    p = N.Params()
    h = ctypes.c_void_p()
    rc = lib.wt_create(ctypes.byref(d), ctypes.byref(p), ctypes.byref(h))
    assert rc == N.WT_E_ARG
    assert b"dims" in lib.wt_last_error()
    assert lib.wt_destroy(None) == N.WT_OK
    assert lib.wt_bytes_for(ctypes.byref(N.Dims(256, 128, 128))) == 256 * 128 * 128 * (19 * 8 + 56)
