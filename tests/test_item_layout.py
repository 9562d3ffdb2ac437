"""The step kernels' bulk item layout, checked on the host from the same header the kernels use
(tests/native/item_layout_check.hip): coverage of every cell exactly once, in-row loads for idle
lanes, uniform rows per item, and the role table's row-to-item map. Grids: the 24 sweep domains,
the benchmark grids, the parity cases' shapes, and edge shapes (rows of one chunk, of 64 cells,
rows one cell past a chunk, odd ny, ny = 2, 2D)."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")


def grids():
    from windtunnel_vis_amd import sweep
    g = set()
    for c in sweep.sweep_cases():
        nx, ny, nz = c["dims"]
        g.add((nx, ny, nz, 2))
    g |= {(512, 256, 256, 2), (256, 128, 128, 2), (256, 128, 128, 1), (512, 128, 1, 2), (512, 128, 1, 1),
          (24, 12, 10, 2), (24, 12, 10, 1), (10, 8, 8, 2), (384, 6, 5, 2), (300, 10, 1, 2), (300, 10, 1, 1),
          (2, 4, 4, 2), (4, 4, 3, 2), (64, 6, 2, 2), (66, 6, 2, 2), (128, 4, 4, 2), (130, 4, 4, 2), (192, 8, 3, 2),
          (194, 8, 3, 2), (200, 7, 5, 2), (200, 2, 5, 2), (160, 3, 3, 2), (97, 6, 4, 1), (160, 80, 1, 2)}
    return sorted(g)


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_item_layout_covers_every_cell_once():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "item_layout_check")
        src = os.path.join(ROOT, "tests", "native", "item_layout_check.hip")
        r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-O1", "-std=c++17", "-Wno-unused-command-line-argument",
                            "-I", os.path.join(ROOT, "include"), src, "-o", exe], stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, text=True)
        assert r.returncode == 0, r.stdout[-3000:]
        args = ["%d,%d,%d,%d" % g for g in grids()]
        r = subprocess.run([exe] + args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
        print(r.stdout)
        assert r.returncode == 0 and r.stdout.startswith("ok %d " % len(args)), r.stdout
