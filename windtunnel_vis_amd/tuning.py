"""Launch autotuner front end: picks cells per thread and the XCD item-run length for a solver
configuration (wt_autotune times every candidate on the device) and keeps the choice in a JSON
cache keyed by lattice, storage, grid and collision, so later runs of the same configuration start
tuned without timing again.
"""
import hashlib
import json
import os

from . import _native as N

_LIB_TAG = None


def lib_tag():
    """Short hash of the built library: a rebuilt kernel invalidates cached choices."""
    global _LIB_TAG
    if _LIB_TAG is None:
        h = hashlib.sha1()
        with open(N.LIB_PATH, "rb") as f:
            h.update(f.read())
        _LIB_TAG = h.hexdigest()[:12]
    return _LIB_TAG


def default_cache():
    return os.environ.get("WT_TUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "windtunnel_vis_amd",
                                                           "tuning.json")


def key_for(solver):
    d = solver.dims
    return "%s/%s/%dx%dx%d/%s/%s" % (solver.lattice.name, solver.precision, d["nx"], d["ny"], d["nz"],
                                     solver.params_.get("collision", "regularized"), lib_tag())


def load_cache(path):
    try:
        with open(path) as f:
            db = json.load(f)
        return db if isinstance(db, dict) else {}
    except (OSError, ValueError):
        return {}


def save_cache(path, db):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp.%d" % os.getpid()
    with open(tmp, "w") as f:
        json.dump(db, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def apply(solver, entry):
    solver.setTuning(N.TUNE_CELLS_PER_THREAD, int(entry["cellsPerThread"]))
    solver.setTuning(N.TUNE_ITEM_CHUNK, int(entry["chunk"]))


def autotune(solver, steps=20, cache=None, refresh=False):
    """Tune `solver` (cached). Returns the entry with `cached` telling whether it was timed now."""
    path = cache or default_cache()
    db = load_cache(path)
    k = key_for(solver)
    if not refresh and k in db:
        apply(solver, db[k])
        return dict(db[k], cached=True)
    r = solver.autotune(steps)
    entry = dict(chunk=r["chunk"], cellsPerThread=r["cellsPerThread"], msPerStep=r["msPerStep"], trials=r["trials"])
    db[k] = entry
    save_cache(path, db)
    return dict(entry, cached=False)
