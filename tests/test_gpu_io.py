"""Solver outputs on the GPU: plot files at the plot times (dt clipped to hit
them, Solver.cpp:197-205), checkpoints every ckptfreq steps and at the end,
and a restart from a checkpoint that continues the run exactly."""
import glob
import os
import re

import numpy as np
import pytest

from helpers import config, init_text

pytestmark = pytest.mark.gpu


def _text(tmp_path, extra=""):
    c = config(dim=2, n=(32, 32, 1), problem="smooth_wave", user=(0.5, 0.25, 0.0), bc="periodic",
               final=0.1, cfl=0.8)
    return init_text(c) + f"io:datadir {tmp_path}\nio:basename sw\nio:plotfreq 0.05\nio:ckptfreq 4\n" + extra


def test_plots_at_plot_times_checkpoints_and_restart(tmp_path):
    import kfvm_amd
    s = kfvm_amd.Solver(init_text=_text(tmp_path))
    s.solve()
    t, _, nts = s.time
    final, pf = float(np.float32(0.1)), float(np.float32(0.05))  # init reals are read as floats
    assert t == final
    U_end = s.state
    d = f"{tmp_path}/R2_NX32_NY32"
    plots = sorted(glob.glob(d + "/plot/sw_*.xmf"))
    times = [float(re.search(r'<Time Value="([^"]+)"', open(p).read()).group(1)) for p in plots]
    print(f"[io] plots {len(plots)} at {times}, steps {nts - 1}", flush=True)
    assert len(plots) == 3 and times[0] == 0.0
    assert times[1] == pf and times[2] == final
    ckpts = sorted(glob.glob(d + "/ckpt/sw_*.h5"))
    steps = [int(os.path.basename(p)[3:10]) for p in ckpts]
    assert steps[0] == 0 and all(k % 4 == 0 for k in steps[1:-1]) and steps[-1] == nts - 1
    # restart from a checkpoint in the middle and run to the end
    mid = [p for p, k in zip(ckpts, steps) if 0 < k < nts - 1][0]
    r = kfvm_amd.Solver(init_text=_text(tmp_path / "r", f"io:restartfile {mid}\n"))
    U_mid, k, tk, dtk, ek = kfvm_amd.read_checkpoint_file(_text(tmp_path), mid, U_end.shape)
    assert r.time[0] == tk and r.time[2] == k + 1
    r.solve()
    diff = np.abs(r.state - U_end).max()
    print(f"[io] restart from step {k}: max |U - U_uninterrupted| = {diff:.3e}", flush=True)
    assert r.time[0] == t and diff == 0.0


def test_kfvm_ex_runs_a_reference_init_file(tmp_path):
    """kfvm.ex <file.init> <problem> on exec/Hydro/Riem_3d_1c/riem.init as the
    reference ships it (tests/golden/riem_3d_1c.init), with only the size, the
    step count and the output folder overridden by later lines: three adaptive
    steps whose final checkpoint equals the oracle's state after three steps of
    the same configuration."""
    import subprocess
    import kfvm_amd
    from helpers import oracle_for
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "kfvm-kokkos_amd", "build", "kfvm.ex")
    with open(os.path.join(root, "tests", "golden", "riem_3d_1c.init")) as f:
        text = f.read()
    text += f"\nmesh:resolution 24 24 24\ntime:maxsteps 4\nio:datadir {tmp_path}\n"
    path = tmp_path / "riem.init"
    path.write_text(text)
    r = subprocess.run([exe, str(path), "riem_3d_1c"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-400:], r.stderr[-400:], flush=True)
    assert r.returncode == 0 and "Time stepping completed" in r.stdout
    ckpts = sorted(glob.glob(f"{tmp_path}/R2_NX24_NY24_NZ24/ckpt/RiemHLLC_*.h5"))
    assert [int(os.path.basename(q)[9:16]) for q in ckpts] == [0, 3]
    full = text + "problem:name riem_3d_1c\n"
    U, step, t, dt, _ = kfvm_amd.read_checkpoint_file(full, ckpts[-1], (5, 24, 24, 24))
    p = kfvm_amd.parse_init(full)
    c = {"equation": "hydro", "riemann": "hllc", "rk": "ssp4_3_2"}
    dx = (p["hi"][0] - p["lo"][0]) / 24
    fw, dw = kfvm_amd.build_weights(3, p["nq"], p["lfac"], dx, dx, dx)
    o = oracle_for(c, p, fw, dw)
    o.set_ic("riem_3d_1c", p["user"], p["seed"])
    fl = np.zeros((2, 24, 24, 24))
    fl[1] = 1.0
    o.flags = fl
    o.prime()
    for _ in range(3):
        assert o.take_step() == 0
    tc = o.time[0]
    d = np.abs(U - o.state).max() / np.abs(o.state).max()
    print(f"[io] kfvm.ex riem_3d_1c 24^3: step {step}, t {t} vs oracle {tc}, max rel diff {d:.3e}", flush=True)
    assert step == 3 and abs(t - tc) <= 1e-12 * tc and d <= 1e-10
