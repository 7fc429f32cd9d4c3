"""Throughput of the KFVM-WENO update on one MI355X, plain vs learned indicator.

Workload: 3D Euler (gamma 1.4) on [0,1]^3 at 256^3 cells, FP64, planar
Riemann problem across x+y+z = 1.5 with (rho,u,v,w,p) = (1,0,0,0,1) /
(0.125,0,0,0,0.1), outflow boundaries, HLLC, KFVM-WENO radius 2 with 2x2
Gauss points per face, SSP(4,3,2) with the adaptive controller, to t = 0.2.
It runs twice: plain (KXRCF selection of the WENO cells) and with the learned
troubled-cell indicator (weights trained on the GPU by
kfvm_amd.train_indicator on other planar problems at 64^3).  Each run starts
at t = 0, takes W untimed steps (default: until t = 0.1, the developed flow),
times K whole RK steps (take_step: reconstruction, selection, WENO, limiter,
Riemann solves, divergence, register update, error norm, controller, and for
the learned run the admissibility check), then finishes to t = 0.2 for the L1
density error against the exact (Sod) solution along the plane normal.  The
state is resident in HBM throughout.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--retrain] [--plain-only]

Prints one JSON line: value = zone-updates of the accepted RK stages per
second in the timed window of the learned run (stages of rejected attempts
cost time but add no zone-update); beside it the plain and learned figures,
each run's wall time from t = 0 to t = 0.2 (time to solution) and L1 error,
the roofline of the dominant kernel from HIP events taken during the timed
window, and the CPU baseline (the oracle on a bounded sample of the same
problem on every core this process may use).
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


def init_text(n, wenothresh, fixeddt=None, nq=2):
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
        f"problem:nquad {nq}",
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


def flop_weno(nq):  # 47,100 at NQ = 2, 78,900 at NQ = 3 (Q = NQ^2 points per face)
    return 2 * 19 * 5 * _SUBSUM + 2 * 8 * 19 * 5 + 2 * 6 * nq * nq * 5 * _SUBSUM


def flop_linear(nq):  # 7,920 at NQ = 2, 17,820 at NQ = 3
    return 2 * 6 * nq * nq * 5 * 33
BYTES_STAGE = 40 * 5                           # read U_s, Uprev, Uhat; write U_s+1, Uhat


def bytes_linear(nq):  # 1,000 at NQ = 2: read U, write the 6 nq^2 face states of 5 variables
    return 40 + 8 * 5 * 6 * nq * nq
# Whole-path FP64 work per zone-update of the same configuration, linear path
# and full WENO (reconstruction, smoothness, WENO faces, Riemann solves at
# ~200 flop-equivalents each, divergence and RK update): the path's FP64
# ceiling at a WENO fraction f is 78.6e12 / (F_LIN + f (F_WENO - F_LIN)).
PATH_F = {2: (12.4e3, 51.9e3), 3: (26.9e3, 87.9e3)}  # (linear path, full WENO) by NQ
# HBM bytes per launch of a kernel at 256^3, NQ = 2, from rocprofv3 FETCH_SIZE /
# WRITE_SIZE passes over this workload (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0: the
# factors a plain copy of known size gives), with the profile they come from;
# None where not measured.
TRAFFIC = {"flux_rk": (23.33e9, "profiles/flux_rk_traffic_riemann3d_256.md"),
           "linear": (20.87e9, "profiles/flux_rk_traffic_riemann3d_256.md")}
# Algorithmic FP64 work per cell of the Riemann / divergence / RK kernel: 3 NQ^2
# face solves per cell at ~200 flop-equivalents each (HLLC: 8 square roots,
# 11 divisions and one power counted as 10 flops each, plus ~70 arithmetic
# operations), the divergence (3 flops per face point and variable) and the
# register update (~8 per variable).
def flop_flux(nq):  # 2,620 at NQ = 2, 5,845 at NQ = 3
    return 3 * nq * nq * 200 + 3 * nq * nq * 5 * 3 + 40
# FP64 operations per cell the counters measure for the two largest kernels at
# 256^3, NQ = 2 (rocprofv3 SQ_INSTS_VALU_FLOPS_FP64 + _TRANS x 64 lanes / cells;
# divisions and square roots count their Newton steps), and the fraction of
# time the VALU is busy (SQ_ACTIVE_INST_VALU x 4 / SIMDs / cycles at the
# kernel's traced duration), from the counter profile.
COUNTERS = {"flux_rk": {"fp64_flop_per_cell": 7754.0, "valu_busy": 0.64},
            "linear": {"fp64_flop_per_cell": 8946.0, "valu_busy": 0.58}}
COUNTERS_SOURCE = "profiles/fp64_counters_riemann3d_256.md"


def _host_cores():
    """Cores this process may run on (the cpuset / affinity mask), and the CPU model."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return n, model


def cpu_baseline(n=64, reps=5):
    """The oracle (C++/OpenMP restatement of the reference's unfused dataflow)
    on n^3 of the same problem with a fixed dt, on every core this process may
    use: per method (KXRCF-sparse and full WENO) one warm-up step (4 stages),
    then the median of `reps` timed steps, in zone-updates per stage per second."""
    import numpy as np
    import kfvm_amd
    from orc import Oracle
    threads, model = _host_cores()
    out = {}
    for label, wt in (("sparse", 1.0), ("full_weno", -1.0)):
        text = init_text(n, wt)
        p = kfvm_amd.parse_init(text)
        f, d = kfvm_amd.build_weights(3, 2, p["lfac"], 1.0 / n, 1.0 / n, 1.0 / n)
        o = Oracle(dim=3, n=(n, n, n), face_w=f, deriv_w=d, nq=2, eq="hydro", rs="hllc",
                   rk="ssp4_3_2", bc="outflow", gamma=p["gamma"], weno_thresh=p["wenothresh"],
                   cfl=p["cfl"], final_time=1e9, atol=p["atol"], rtol=p["rtol"],
                   fixed_dt=0.0025, threads=threads)
        o.set_ic("riemann3d_planar", (1.5,))
        o.prime()
        o.take_step()  # warm-up: 4 stages
        times = []
        for _ in range(reps):
            e0 = o.rhs_evals
            t0 = time.perf_counter()
            o.take_step()
            times.append((time.perf_counter() - t0) / (o.rhs_evals - e0))
        out[label] = n ** 3 / float(np.median(times)) / 1e6
    return {"value": out["sparse"], "unit": "Mzone-updates/s", "cores": threads, "kind": "port",
            "full_weno_value": out["full_weno"], "cpu": model,
            "sample": f"oracle (C++/OpenMP, the reference's unfused dataflow) on {n}^3 of the same problem, "
                      f"fixed dt, {threads} threads: 1 warm-up step, then the median of {reps} steps per method; "
                      f"value = KXRCF-sparse, full_weno_value = full WENO"}


def cpu_workloads(big=256, big_steps=2, warm_stages=3, log=None):
    """The oracle on the two CPU workloads of the path's measurement plan, on
    every core this process may use, for KXRCF-sparse and full WENO:
      * 64^3 from t = 0 to t = 0.2 with the adaptive controller (wall time,
        steps, zone-updates per stage per second);
      * big^3 with a fixed dt: warm_stages untimed RHS evaluations (the FSAL
        prime and the first stages), then big_steps timed steps (their median).
    No GPU is used."""
    import numpy as np
    import kfvm_amd
    from orc import Oracle
    threads, model = _host_cores()
    log = log or (lambda m: print(m, file=sys.stderr, flush=True))

    def make(n, wt, fixed):
        text = init_text(n, wt)
        p = kfvm_amd.parse_init(text)
        f, d = kfvm_amd.build_weights(3, 2, p["lfac"], 1.0 / n, 1.0 / n, 1.0 / n)
        o = Oracle(dim=3, n=(n, n, n), face_w=f, deriv_w=d, nq=2, eq="hydro", rs="hllc",
                   rk="ssp4_3_2", bc="outflow", gamma=p["gamma"], weno_thresh=p["wenothresh"],
                   cfl=p["cfl"], final_time=0.2, atol=p["atol"], rtol=p["rtol"],
                   fixed_dt=fixed, threads=threads, max_steps=100000)
        o.set_ic("riemann3d_planar", (1.5,))
        return o

    out = {"cores": threads, "cpu": model}
    for label, wt in (("sparse", 1.0), ("full_weno", -1.0)):
        o = make(64, wt, -1.0)
        t0 = time.perf_counter()
        o.prime()
        steps = 0
        while o.time[0] < 0.2 - 1e-12:
            o.take_step()
            steps += 1
        el = time.perf_counter() - t0
        out[f"n64_to_t02_{label}"] = {"seconds": el, "steps": steps, "rhs_evals": o.rhs_evals,
                                      "mzone_updates_per_s": 64 ** 3 * o.rhs_evals / el / 1e6}
        log(f"[cpu] 64^3 {label} to t = 0.2: {el:.1f} s, {steps} steps")
        o = make(big, wt, 0.0025)
        o.prime()
        while o.rhs_evals < warm_stages:
            o.eval_rhs(0)
        times = []
        for k in range(big_steps):
            e0 = o.rhs_evals
            t0 = time.perf_counter()
            o.take_step()
            times.append((time.perf_counter() - t0) / (o.rhs_evals - e0))
            log(f"[cpu] {big}^3 {label} step {k + 1}: {times[-1]:.2f} s per stage")
        out[f"n{big}_{label}"] = {"s_per_stage_median": float(np.median(times)), "steps": big_steps,
                                  "mzone_updates_per_s": big ** 3 / float(np.median(times)) / 1e6}
    return out


def _mode_run(n, weights, a, profile):
    """One run of the planar Riemann problem from t = 0 to t = 0.2: untimed
    warm-up steps, a timed window of K whole RK steps (take_step: stages,
    error norm, controller, admissibility check), then the rest of the run.
    Kernel times are HIP events on the solver's stream during the timed
    window; the whole run's wall time is the time to solution."""
    import numpy as np
    import kfvm_amd
    from kfvm_amd.exact import planar_riemann_l1
    s = kfvm_amd.Solver(init_text=init_text(n, a.wenothresh, nq=a.nq))
    if weights is not None:
        s.set_indicator(weights, a.threshold)
    s.sync()
    t_run = time.perf_counter()
    nwarm = 0
    if a.warmup >= 0:
        for _ in range(a.warmup):
            s.take_step()
            nwarm += 1
    else:
        while s.time[0] < 0.1 and not s.time[0] >= 0.2:
            s.take_step()
            nwarm += 1
    s.sync()
    c0 = s.counters
    ind0 = s.indicator_counters if weights is not None else None
    if profile:
        s.profile(True)
    t0 = time.perf_counter()
    nweno = 0
    accepted = 0
    for _ in range(a.steps):
        st = s.take_step()
        accepted += st == 0
        nweno += s.counters["last_nweno"]
    s.sync()
    el = time.perf_counter() - t0
    c1 = s.counters
    stages = c1["rhs_evals"] - c0["rhs_evals"]
    rejected = (c1["rejects_tol"] - c0["rejects_tol"]) + (c1["rejects_unphys"] - c0["rejects_unphys"])
    kern = {}
    if profile:
        s.profile(False)
        for name in ("cell_props", "linear", "select", "weno", "weno_post", "pp", "flux_rk", "stage_pre", "nad"):
            ms, cnt = s.profile_get(name)
            if cnt:
                kern[name] = {"ms_total": ms, "launches": cnt, "ms_avg": ms / cnt}
    ind1 = s.indicator_counters if weights is not None else None
    while s.time[0] < 0.2 - 1e-12 and not a.no_l1:
        if s.take_step() not in (0, 1, 2):
            break
    s.sync()
    t_solution = time.perf_counter() - t_run
    l1 = planar_riemann_l1(s.state[0], s.time[0])
    rk_stages = 4  # SSP(4,3,2)
    out = {"value": n ** 3 * rk_stages * accepted / el / 1e6,
           "value_all_stages": n ** 3 * stages / el / 1e6,
           "ms_per_step": el / a.steps * 1e3, "warmup_steps": nwarm, "stages": stages,
           "accepted_steps": accepted, "rejected_attempts": rejected, "seconds": el,
           "weno_fraction": nweno / a.steps / n ** 3, "l1_rho_t0.2": l1, "t_final": s.time[0],
           "steps_total": s.time[2] - 1, "rejects_total": s.counters["rejects_tol"] + s.counters["rejects_unphys"],
           "seconds_to_solution": t_solution}
    if weights is not None:
        out.update(s.indicator_counters)
        out["nad_cells_per_step_timed"] = (ind1["nad_cells"] - ind0["nad_cells"]) / a.steps
    s.close()
    return out, kern


def _roofline(kern, head, ncell, n, nq):
    """Roofline of the kernel with the largest profiled time, plus the whole
    path against its FP64 ceiling at the measured WENO fraction.  The
    reconstruction, WENO and flux kernels do more FP64 work per byte of
    algorithmic traffic than the ridge point (78.6 TFLOP/s / 8 TB/s = 9.8
    flop/B; the flux kernel ~13), so their roof is FP64 compute ("mfma": the
    FP64 peak, which the vector FMA pipe shares); the HBM figures are given
    beside it."""
    dom = max(kern, key=lambda k: kern[k]["ms_total"])
    d = kern[dom]
    t = d["ms_avg"] * 1e-3
    if dom == "weno":
        alg = flop_weno(nq) * head["weno_fraction"] * ncell
    elif dom == "linear":
        alg = flop_linear(nq) * ncell
    else:
        alg = flop_flux(nq) * ncell
    roof = {"bound": "mfma", "achieved": alg / t / 1e12, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "algorithmic_flop_per_launch": alg}
    roof["frac"] = roof["achieved"] / roof["peak"]
    tr = TRAFFIC.get(dom) if n == 256 and nq == 2 else None
    roof["traffic"] = tr[0] if tr else None
    roof["traffic_source"] = tr[1] if tr else None
    roof["kernel"] = dom
    # the kernel's algorithmic traffic against HBM: the RK stage's 200 B per
    # cell for the flux kernel, U in and the face states out for the linear pass
    nbytes = bytes_linear(nq) if dom == "linear" else BYTES_STAGE
    roof["hbm"] = {"achieved": nbytes * ncell / t / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                   "frac": nbytes * ncell / t / 1e9 / PEAK_HBM_GBS, "bytes_per_cell": nbytes,
                   "traffic_GBps": (tr[0] / t / 1e9) if tr else None}
    cn = COUNTERS.get(dom) if n == 256 and nq == 2 else None
    if cn:
        tf = cn["fp64_flop_per_cell"] * ncell / t / 1e12
        roof["fp64_counters"] = {"tflops": tf, "frac": tf / PEAK_FP64_TFLOPS, "valu_busy": cn["valu_busy"],
                                 "source": COUNTERS_SOURCE}
    # the whole path against its FP64 ceiling at the measured WENO fraction
    f_lin, f_weno = PATH_F[nq]
    f_alg = f_lin + head["weno_fraction"] * (f_weno - f_lin)
    ceil = PEAK_FP64_TFLOPS * 1e12 / f_alg / 1e6
    roof["path"] = {"bound": "fp64", "flop_per_zone_update": f_alg, "ceiling_mzone_updates_per_s": ceil,
                    "frac": head["value"] / ceil}
    return roof


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed RK steps per method")
    ap.add_argument("--warmup", type=int, default=-1,
                    help="untimed RK steps first; -1: until t = 0.1 (the developed flow)")
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--nq", type=int, default=2, choices=(2, 3), help="quadrature points per face direction")
    ap.add_argument("--wenothresh", type=float, default=1.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-only", action="store_true")
    ap.add_argument("--learned-only", action="store_true")
    ap.add_argument("--no-l1", action="store_true", help="stop after the timed window (profiling runs)")
    ap.add_argument("--retrain", action="store_true", help="train the indicator on the GPU first")
    ap.add_argument("--threshold", type=float, default=0.0,
                    help="logit above which the learned indicator selects WENO")
    ap.add_argument("--fused", action="store_true",
                    help="fused reconstruction + flux path (the flux kernel recomputes the linear states)")
    ap.add_argument("--cpu-workloads", type=int, default=0, metavar="N",
                    help="only time the CPU oracle: 64^3 to t = 0.2 and N^3 for 2 fixed-dt steps (no GPU)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (5 launches), the "
                         "reference kernel for scaling FETCH_SIZE / WRITE_SIZE counters")
    a = ap.parse_args()
    if a.cpu_workloads:
        print(json.dumps(cpu_workloads(big=a.cpu_workloads)), flush=True)
        return
    if a.calibrate_copy:
        import torch
        nn = a.calibrate_copy * (1 << 20) // 8
        src = torch.rand(nn, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nn * 8, "copy_bytes_written": nn * 8, "launches": 5}))
        return
    if a.fused:
        os.environ["KFVM_FUSED"] = "1"
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the domain decomposition over GPUs is design-only)", file=sys.stderr)
        sys.exit(2)

    import kfvm_amd
    from kfvm_amd import train_indicator

    n = a.size
    weights = None
    if not a.plain_only:
        if a.retrain:
            import tempfile
            # the retrained weights go to a scratch directory, not over the committed ones
            weights, _ = train_indicator.train("euler3d", out_dir=tempfile.mkdtemp(prefix="kfvm_weights_"),
                                               log=lambda m: print(m, flush=True))
        else:
            weights = train_indicator.load("euler3d")
    plain, kp = None, {}
    if not a.learned_only or weights is None:
        plain, kp = _mode_run(n, None, a, weights is None)
        print(f"[bench] plain: {plain['value']:.1f} Mzone-updates/s (accepted stages), WENO "
              f"{plain['weno_fraction']:.3f}, L1 {plain['l1_rho_t0.2']:.5e}, {plain['steps_total']} steps "
              f"in {plain['seconds_to_solution']:.2f} s to t = 0.2", flush=True)
    learned, kl = (None, {})
    if weights is not None:
        learned, kl = _mode_run(n, weights, a, True)
        print(f"[bench] learned: {learned['value']:.1f} Mzone-updates/s (accepted stages), WENO "
              f"{learned['weno_fraction']:.3f}, L1 {learned['l1_rho_t0.2']:.5e}, {learned['steps_total']} steps "
              f"in {learned['seconds_to_solution']:.2f} s to t = 0.2, redone steps {learned['nad_fallbacks']}",
              flush=True)
    head = learned if learned is not None else plain
    kern = kl if learned is not None else kp
    ncell = n ** 3
    roof = _roofline(kern, head, ncell, n, a.nq) if kern else None
    cpu = None if a.no_cpu else cpu_baseline()
    method = "learned indicator" if learned is not None else "plain (KXRCF)"
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "Mzone-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": head["warmup_steps"],
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D Euler planar Riemann {n}^3 (x+y+z=1.5, (1,0,0,0,1)/(0.125,0,0,0,0.1)), "
                               f"HLLC, KFVM-WENO r=2 NQ={a.nq}, SSP(4,3,2) adaptive, to t=0.2; value: {method}",
                   "cells": ncell, "step": "one RK step (4 stages); value counts zone-updates per stage"},
        "plain": plain,
        "learned": learned,
        "speedup_learned_vs_plain": (learned["value"] / plain["value"]) if learned and plain else None,
        "time_to_solution_ratio_plain_over_learned": (plain["seconds_to_solution"] / learned["seconds_to_solution"])
        if learned and plain else None,
        "l1_ratio_learned_vs_plain": (learned["l1_rho_t0.2"] / plain["l1_rho_t0.2"]) if learned and plain else None,
        "roofline": roof,
        "kernels_ms_avg": {k: v["ms_avg"] for k, v in kern.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
