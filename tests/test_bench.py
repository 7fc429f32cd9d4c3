"""bench.py's host-side contract, on the CPU: one GPU only, and the
algorithmic work figures its roofline uses."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr) or "1 GPU" in (r.stdout + r.stderr)


def test_work_figures_match_the_stencil_counts():
    import bench
    # 3D, r = 2: C = 33 stencil cells, substencil columns 33 + 7 + 6*11 = 106,
    # I = 19 smoothness rows, F = 6 faces, NV = 5
    assert bench.flop_linear(2) == 2 * 6 * 4 * 5 * 33 == 7920
    assert bench.flop_linear(3) == 17820
    assert bench.flop_weno(2) == 2 * 19 * 5 * 106 + 2 * 8 * 19 * 5 + 2 * 6 * 4 * 5 * 106 == 47100
    assert bench.flop_weno(3) == 78900
    assert bench.BYTES_STAGE == 200
    lin, weno = bench.PATH_F[2]
    assert lin < weno
