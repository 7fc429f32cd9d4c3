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


def test_roofline_of_the_dominant_kernel():
    import bench
    head = {"weno_fraction": 0.05, "value": 900.0}
    kern = {"linear": {"ms_total": 10.0, "ms_avg": 1.0},
            "flux_rk": {"ms_total": 20.0, "ms_avg": 2.0}}
    r = bench._roofline(kern, head, 256 ** 3, 256, 2)
    # the flux kernel's roof is FP64 compute: 12 face solves of ~200 flops, the
    # divergence and the register update per cell, over 2 ms
    assert r["kernel"] == "flux_rk" and r["bound"] == "mfma" and r["unit"] == "TFLOP/s"
    assert bench.flop_flux(2) == 12 * 200 + 12 * 5 * 3 + 40
    assert abs(r["achieved"] - bench.flop_flux(2) * 256 ** 3 / 2e-3 / 1e12) < 1e-12
    assert abs(r["frac"] - r["achieved"] / bench.PEAK_FP64_TFLOPS) < 1e-15
    # beside it: 200 B per cell of register traffic over 2 ms, and the counters
    assert abs(r["hbm"]["achieved"] - 200 * 256 ** 3 / 2e-3 / 1e9) < 1e-9
    assert abs(r["fp64_counters"]["tflops"] - 7754.0 * 256 ** 3 / 2e-3 / 1e12) < 1e-9
    assert r["traffic"] == bench.TRAFFIC["flux_rk"][0] and r["traffic_source"].startswith("profiles/")
    lin, weno = bench.PATH_F[2]
    assert abs(r["path"]["flop_per_zone_update"] - (lin + 0.05 * (weno - lin))) < 1e-9
    # another size: no counter figure for the traffic
    assert bench._roofline(kern, head, 128 ** 3, 128, 2)["traffic"] is None


def test_roofline_when_the_linear_pass_dominates():
    """The linear pass and the flux kernel take about the same time; when the
    linear pass takes more, its own algorithmic flops (the 24-point
    contraction of five variables over 33 cells) and bytes (U in, the face
    states out: 40 + 960 B per cell at NQ=2) are priced."""
    import bench
    head = {"weno_fraction": 0.04, "value": 1000.0}
    kern = {"linear": {"ms_total": 21.0, "ms_avg": 2.1},
            "flux_rk": {"ms_total": 20.0, "ms_avg": 2.0}}
    r = bench._roofline(kern, head, 256 ** 3, 256, 2)
    assert r["kernel"] == "linear"
    assert bench.bytes_linear(2) == 40 + 24 * 5 * 8 == 1000
    assert abs(r["achieved"] - 7920 * 256 ** 3 / 2.1e-3 / 1e12) < 1e-12
    assert abs(r["hbm"]["achieved"] - 1000 * 256 ** 3 / 2.1e-3 / 1e9) < 1e-9
    assert r["traffic"] == bench.TRAFFIC["linear"][0]
    assert abs(r["fp64_counters"]["tflops"] - bench.COUNTERS["linear"]["fp64_flop_per_cell"] * 256 ** 3
               / 2.1e-3 / 1e12) < 1e-9
