"""The seeded sweep generator (host logic only)."""
from windtunnel_vis_amd import sweep


def test_splitmix64_known_values():
    # splitmix64 reference outputs for seed 0 (Steele, Lea, Flood 2014; the java.util.SplittableRandom mix)
    r = sweep.SplitMix64(0)
    assert r.next() == 0xE220A8397B1DCDAF
    assert r.next() == 0x6E789E6AA1B965F4
    assert r.next() == 0x06C45D188009454F


def test_cases_are_deterministic_and_in_range():
    a = sweep.sweep_cases()
    b = sweep.sweep_cases()
    assert a == b and len(a) == 24
    for c in a:
        nx, ny, nz = c["dims"]
        assert 128 <= nx <= 320 and 64 <= ny <= 160 and 64 <= nz <= 160
        assert nx % 16 == 0 and ny % 16 == 0 and nz % 16 == 0
        assert -15 <= c["yaw"] <= 15 and 2 <= c["ride"] <= 6
        assert c["kind"] in ("sedan", "suv")
    assert len({c["dims"] for c in a}) > 5  # uneven grids


def test_case_body_fits_and_rides():
    for c in sweep.sweep_cases()[:6]:
        body = sweep.case_body(c)
        nx, ny, nz = c["dims"]
        assert abs(body.lowest_y()) < 1e-9  # wheels on the floor
        assert abs(body.length - nx / 3.0) < 1e-9
