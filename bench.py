"""Throughput of the KFVM-WENO per-stage update on one MI355X.

Workload: 3D Euler (gamma 1.4) on [0,1]^3 at 256^3 cells, FP64, planar
Riemann problem across x+y+z = 1.5 with (rho,u,v,w,p) = (1,0,0,0,1) /
(0.125,0,0,0,0.1), outflow boundaries, HLLC, KFVM-WENO radius 2 with 2x2
Gauss points per face, SSP(4,3,2) stages, sparse (KXRCF) WENO with threshold
1.0 -- the plain method.  One bench "step" is one RK stage: reconstruction,
selection, WENO, positivity limiter, Riemann solves, flux divergence,
register update (and the error norm on the last stage of each RK step), over
every cell, with the state resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line (metric, value in Mzone-updates/s, roofline of the
dominant kernel from HIP events, CPU baseline from the oracle on a bounded
sample of the same problem).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "kfvm-kokkos_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "Mzone-updates/s (3D Euler, KFVM-WENO r=2) plain vs learned-indicator, equal L1"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def init_text(n, wenothresh, fixeddt=None):
    lines = [
        f"mesh:resolution {n} {n} {n}",
        "mesh:extent 0.0 1.0 0.0 1.0 0.0 1.0",
        "stencil:lfac 5",
        "time:final 0.2",
        "time:cfl 1.25",
        "time:atol 1.e-3",
        "time:rtol 1.e-3",
        "eos:gamma 1.4",
        f"eos:wenothresh {wenothresh}",
        "problem:name riemann3d_planar",
        "problem:dim 3",
        "problem:nquad 2",
        "problem:equation hydro",
        "problem:riemann hllc",
        "problem:rk ssp4_3_2",
        "problem:bc outflow",
        "user:p0 1.5",
    ]
    if fixeddt:
        lines.append(f"time:fixeddt {fixeddt!r}")
    return "\n".join(lines) + "\n"


# Algorithmic FP64 work per cell of each kernel family, 3D, radius 2, NQ = 2
# (FMA = 2 flops), from the stencil sizes C=33, Cc=7, Cb=11, B=6, I=19,
# F=6 faces, Q=4 points, NV=5.
_SUBSUM = 33 + 7 + 6 * 11                      # 106 stencil columns over all substencils
FLOP_WENO = 2 * 19 * 5 * _SUBSUM + 2 * 8 * 19 * 5 + 2 * 6 * 4 * 5 * _SUBSUM  # 47,100
FLOP_LINEAR = 2 * 6 * 4 * 5 * 33               # 7,920
BYTES_STAGE = 40 * 5                           # read U_s, Uprev, Uhat; write U_s+1, Uhat


def cpu_baseline(seconds_target=15.0):
    """The oracle (C++/OpenMP restatement, reference dataflow) on 48^3 of the
    same problem, fixed dt, timed over whole RK steps."""
    import numpy as np
    import kfvm_amd
    from orc import Oracle
    threads = min(os.cpu_count() or 1, 16)
    n = 48
    text = init_text(n, 1.0)
    p = kfvm_amd.parse_init(text)
    f, d = kfvm_amd.build_weights(3, 2, p["lfac"], 1.0 / n, 1.0 / n, 1.0 / n)
    o = Oracle(dim=3, n=(n, n, n), face_w=f, deriv_w=d, nq=2, eq="hydro", rs="hllc",
               rk="ssp4_3_2", bc="outflow", gamma=p["gamma"], weno_thresh=p["wenothresh"],
               cfl=p["cfl"], final_time=1e9, atol=p["atol"], rtol=p["rtol"],
               fixed_dt=0.002, threads=threads)
    o.set_ic("riemann3d_planar", (1.5,))
    o.prime()
    o.take_step()  # warm-up step
    e0 = o.rhs_evals
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < seconds_target and steps < 50:
        o.take_step()
        steps += 1
    el = time.perf_counter() - t0
    stages = o.rhs_evals - e0
    return {"value": n ** 3 * stages / el / 1e6, "unit": "Mzone-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle (C++/OpenMP, reference unfused dataflow) on {n}^3 of the same "
                      f"problem, {stages} RK stages ({steps} fixed-dt steps) in {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--wenothresh", type=float, default=1.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-stages", type=int, default=4)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (5 launches), the "
                         "reference kernel for scaling FETCH_SIZE / WRITE_SIZE counters")
    a = ap.parse_args()
    if a.calibrate_copy:
        import torch
        n = a.calibrate_copy * (1 << 20) // 8
        src = torch.rand(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 5}))
        return
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the domain decomposition over GPUs is design-only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import kfvm_amd

    n = a.size
    t_setup = time.perf_counter()
    s = kfvm_amd.Solver(init_text=init_text(n, a.wenothresh))
    _, dt0, _ = s.time
    dt = 100.0 * dt0  # cfl * dmin / vmax of the initial state
    print(f"[bench] setup {time.perf_counter() - t_setup:.1f} s, {n}^3, dt {dt:.4g}", flush=True)

    s.advance_stages(a.warmup, dt)
    s.sync()
    t0 = time.perf_counter()
    s.advance_stages(a.steps, dt)
    s.sync()
    el = time.perf_counter() - t0
    ncell = n ** 3
    value = ncell * a.steps / el / 1e6
    print(f"[bench] {a.steps} stages in {el:.3f} s -> {value:.1f} Mzone-updates/s", flush=True)

    # per-kernel device time (HIP events on the solver's stream), separate pass
    s.profile(True)
    flagged = []
    for _ in range(a.profile_stages):
        s.advance_stages(1, dt)
        if a.wenothresh > 0:
            flagged.append(float(s.flags[0].sum()))
    s.profile(False)
    kern = {}
    for name in ("cell_props", "linear", "select", "weno", "weno_post", "pp", "flux_rk",
                 "stage_pre"):
        ms, cnt = s.profile_get(name)
        if cnt:
            kern[name] = {"ms_total": ms, "launches": cnt, "ms_avg": ms / cnt}
    dom = max(kern, key=lambda k: kern[k]["ms_total"])
    d = kern[dom]
    if dom == "weno":
        cells = (sum(flagged) / len(flagged)) if flagged else ncell
        flops = FLOP_WENO * cells
        roof = {"bound": "mfma", "achieved": flops / (d["ms_avg"] * 1e-3) / 1e12,
                "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s"}
    elif dom == "linear":
        flops = FLOP_LINEAR * ncell
        roof = {"bound": "mfma", "achieved": flops / (d["ms_avg"] * 1e-3) / 1e12,
                "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s"}
    else:
        # flux/RK kernel: HBM roof on its algorithmic register traffic
        byts = BYTES_STAGE * ncell
        roof = {"bound": "hbm", "achieved": byts / (d["ms_avg"] * 1e-3) / 1e9,
                "peak": PEAK_HBM_GBS, "unit": "GB/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = None
    roof["kernel"] = dom
    frac_flagged = (sum(flagged) / len(flagged) / ncell) if flagged else 1.0

    cpu = None if a.no_cpu else cpu_baseline()
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "Mzone-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": el / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D Euler planar Riemann {n}^3, HLLC, KFVM-WENO r=2 NQ=2, "
                               f"SSP(4,3,2), plain (KXRCF sparse WENO, threshold {a.wenothresh})",
                   "cells": ncell, "weno_fraction": frac_flagged, "step": "one RK stage"},
        "roofline": roof,
        "kernels_ms_avg": {k: v["ms_avg"] for k, v in kern.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
