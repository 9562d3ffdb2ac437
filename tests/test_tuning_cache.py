"""The autotuner's JSON cache (host logic only, no device)."""
import json

from windtunnel_vis_amd import tuning


class FakeSolver:
    def __init__(self):
        from windtunnel_vis_amd.solver import D3Q19
        self.dims = dict(nx=64, ny=32, nz=32)
        self.lattice = D3Q19
        self.precision = "f16"
        self.params_ = {"collision": "regularized"}
        self.set = {}
        self.timed = 0

    def setTuning(self, k, v):
        self.set[k] = v

    def autotune(self, steps):
        self.timed += 1
        return dict(chunk=8, cellsPerThread=2, msPerStep=0.5,
                    trials=[dict(chunk=0, cellsPerThread=2, msPerStep=0.7), dict(chunk=8, cellsPerThread=2, msPerStep=0.5)])


def test_cache_round_trip(tmp_path, monkeypatch):
    monkeypatch.setattr(tuning, "lib_tag", lambda: "testlib")
    path = str(tmp_path / "c" / "tune.json")
    s = FakeSolver()
    r = tuning.autotune(s, cache=path)
    assert not r["cached"] and s.timed == 1
    db = json.load(open(path))
    assert db["D3Q19/f16/64x32x32/regularized/testlib"]["chunk"] == 8
    s2 = FakeSolver()
    r2 = tuning.autotune(s2, cache=path)
    assert r2["cached"] and s2.timed == 0 and s2.set == {1: 2, 0: 8}
    r3 = tuning.autotune(s2, cache=path, refresh=True)
    assert not r3["cached"] and s2.timed == 1


def test_corrupt_cache_is_ignored(tmp_path, monkeypatch):
    monkeypatch.setattr(tuning, "lib_tag", lambda: "testlib")
    path = tmp_path / "tune.json"
    path.write_text("{not json")
    s = FakeSolver()
    r = tuning.autotune(s, cache=str(path))
    assert not r["cached"] and s.timed == 1
