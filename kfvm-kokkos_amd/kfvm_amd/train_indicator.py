"""Train the learned troubled-cell indicator on the GPU.

Labels come from the solver itself: snapshots of plain (KXRCF) runs are
reconstructed once linearly and once with full WENO (+ limiter), and a cell is
labelled 1 where the two differ by more than tau of its stencil's range
(kfvm_indicator_collect).  The network (2C features -> 32 -> 32 -> logit,
bfloat16 weights) is trained by the library's own forward / backward / Adam
kernels.  The weights are written as a .npy of bfloat16 bit patterns (uint16)
plus a .json of the recipe and the training statistics.

    python -m kfvm_amd.train_indicator euler3d [--out DIR]
    python -m kfvm_amd.train_indicator shear2d [--out DIR]
    python -m kfvm_amd.train_indicator mhd3d [--out DIR] [--evaluate [--n 128]]
"""
import argparse
import json
import os
import time

import numpy as np

import kfvm_amd

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "weights")


def init_text(dim, n, problem, user=(), seed=0, bc="outflow", final=1.0, gamma=1.4, wenothresh=1.0,
              nq=2, cfl=None, equation="hydro", riemann="hllc"):
    lines = ["mesh:resolution " + " ".join(str(x) for x in n[:dim]),
             "mesh:extent " + " ".join("0.0 1.0" for _ in range(dim)),
             "stencil:lfac 5", f"time:final {final!r}", "time:atol 1.e-3", "time:rtol 1.e-3",
             f"eos:gamma {gamma!r}", f"eos:wenothresh {wenothresh!r}", f"problem:name {problem}",
             f"problem:dim {dim}", f"problem:nquad {nq}", f"problem:equation {equation}",
             f"problem:riemann {riemann}", "problem:rk ssp4_3_2", f"problem:bc {bc}", f"user:seed {seed}"]
    if cfl is not None:
        lines.append(f"time:cfl {cfl!r}")
    for i, u in enumerate(user):
        lines.append(f"user:p{i} {u!r}")
    return "\n".join(lines) + "\n"


_MHD = dict(equation="mhd_glm", riemann="llf", gamma=5.0 / 3.0)

# Training problems: never the evaluation problem itself at its own resolution.
RECIPES = {
    # 3D Euler planar Riemann problems (the evaluation is the x+y+z = 1.5 plane at
    # 256^3): other plane offsets at 64^3, snapshots every 4th step of 24
    "euler3d": dict(dim=3, tau=0.15, iters=2000, batch=8192, lr=2e-3, seed=1,
                    runs=[dict(n=(64, 64, 64), problem="riemann3d_planar", user=(off,), bc="outflow",
                               steps=24, every=4) for off in (1.2, 1.4, 1.65, 1.8)]),
    # 2D shear layer (the evaluation is 1024^2): 256^2 and 512^2, seeds 0-4,
    # snapshots every 0.1 up to t = 1
    "shear2d": dict(dim=2, tau=0.15, iters=2000, batch=8192, lr=2e-3, seed=1,
                    runs=[dict(n=(m, m), problem="shear2d", user=(), bc="periodic", seed=sd, t_end=1.0,
                               dt_snap=0.1) for m in (256, 512) for sd in range(5)]),
    # 3D GLM-MHD (the evaluation is the periodic vortex at 256^3 to t = 0.5):
    # plain runs of magnetised blasts at 64^3 and of the vortex at 64^3 and
    # 128^3, snapshots every 0.05 (blasts) / 0.1 (vortex).  A lower label threshold
    # and logit threshold than Euler's: the flow is smooth but steep almost
    # everywhere.  The admissibility check is off: it flags about 6% of the
    # cells in every step whatever the selection (growing smooth extrema), so
    # its redo only costs time.
    "mhd3d": dict(dim=3, tau=0.02, threshold=-2.0, nad=False, iters=2000, batch=8192, lr=2e-3, seed=1,
                  runs=[dict(n=(64, 64, 64), problem="mhd_blast", user=u, bc="periodic", t_end=0.2,
                             dt_snap=0.05, **_MHD) for u in ((0.1, 0.5, 0.0), (0.15, 0.0, 1.0))]
                  + [dict(n=(m, m, m), problem="mhd_vortex3d", user=(), bc="periodic", t_end=0.5,
                          dt_snap=0.1, **_MHD) for m in (64, 128)]),
}


def collect(ind, run, tau, dim):
    text = init_text(dim, run["n"], run["problem"], run.get("user", ()), run.get("seed", 0), run["bc"],
                     final=run.get("t_end", 1.0), gamma=run.get("gamma", 1.4),
                     equation=run.get("equation", "hydro"), riemann=run.get("riemann", "hllc"))
    s = kfvm_amd.Solver(init_text=text)
    if "steps" in run:
        for k in range(run["steps"]):
            s.take_step()
            if k % run["every"] == run["every"] - 1:
                ind.collect(s, tau)
    else:
        nxt = run["dt_snap"]
        while True:
            st = s.take_step()
            t = s.time[0]
            if t >= nxt - 1e-12:
                ind.collect(s, tau)
                nxt += run["dt_snap"]
            if t >= run["t_end"] - 1e-12 or st not in (0, 1, 2):
                break
    s.close()


def train(name, out_dir=WEIGHTS, log=print):
    r = RECIPES[name]
    dim = r["dim"]
    t0 = time.time()
    ind = kfvm_amd.Indicator(dim)
    for run in r["runs"]:
        collect(ind, run, r["tau"], dim)
    total = ind.nsamples
    X, Y = ind.samples(0, total)
    pos = float(Y.mean())
    log(f"[train] {name}: {total} samples, {pos:.4f} labelled ({time.time() - t0:.1f} s)")
    ind.init(r["seed"])
    pw = (1.0 - pos) / max(pos, 1e-6)
    loss = ind.train(r["iters"], r["batch"], r["lr"], pos_weight=pw, seed=r["seed"])
    W = ind.weights
    P = kfvm_amd.bf16_to_float(W)
    sub = slice(None, None, max(1, total // 500000))
    z = ind.logits(P, X[sub])
    y = Y[sub] > 0.5
    flagged = z > 0.0
    stats = {"samples": int(total), "positive_fraction": pos, "pos_weight": pw, "last_batch_loss": loss,
             "recall": float((flagged & y).sum() / max(1, y.sum())), "flagged_fraction": float(flagged.mean()),
             "precision": float((flagged & y).sum() / max(1, flagged.sum())),
             "train_seconds": time.time() - t0}
    log(f"[train] {name}: loss {loss:.4f}, recall {stats['recall']:.4f}, flagged {stats['flagged_fraction']:.4f}")
    os.makedirs(out_dir, exist_ok=True)
    np.save(os.path.join(out_dir, f"{name}.npy"), W)
    meta = {k: v for k, v in r.items() if k != "runs"}
    meta["runs"] = [{k: (list(v) if isinstance(v, tuple) else v) for k, v in run.items()} for run in r["runs"]]
    meta["stats"] = stats
    with open(os.path.join(out_dir, f"{name}.json"), "w") as f:
        json.dump(meta, f, indent=1)
    ind.close()
    return W, stats


EVALUATIONS = {
    # 2D shear layer at 1024^2, seed 0, periodic, to t = 1
    "shear2d": dict(dim=2, n=(1024, 1024), problem="shear2d", user=(), bc="periodic", seed=0, t_end=1.0),
    # 3D GLM-MHD periodic vortex at 256^3, LLF, to t = 0.5
    "mhd3d": dict(dim=3, n=(256, 256, 256), problem="mhd_vortex3d", user=(), bc="periodic", seed=0,
                  t_end=0.5, **_MHD),
}


def evaluate(name, weights=None, n=None, log=print, modes=("full", "kxrcf", "learned"), t_end=None):
    """Full WENO (the reference's plain WENO), KXRCF-sparse and learned runs of
    the recipe's evaluation problem to its end time: wall time, zone-updates
    per second, mean WENO fraction, and each run's L1 distance (mean over
    cells and variables) from the full-WENO solution (from the KXRCF one when
    the full-WENO run is left out)."""
    import kfvm_amd as k
    e = EVALUATIONS[name]
    r = RECIPES[name]
    W = load(name) if weights is None else weights
    size = n or e["n"]
    t_end = t_end or e["t_end"]
    res = {}
    states = {}
    for mode in modes:
        text = init_text(e["dim"], size, e["problem"], e["user"], e["seed"], e["bc"], final=t_end,
                         wenothresh=-1.0 if mode == "full" else 1.0, gamma=e.get("gamma", 1.4),
                         equation=e.get("equation", "hydro"), riemann=e.get("riemann", "hllc"))
        s = k.Solver(init_text=text)
        if mode == "learned":
            s.set_indicator(W, r.get("threshold", 0.0))
            s.set_nad(r.get("nad", True))
        ncell = int(np.prod(size))
        fr = []
        s.sync()
        t0 = time.time()
        t_log = t0
        while s.time[0] < t_end - 1e-12:
            st = s.take_step()
            if st not in (0, 1, 2):
                break
            fr.append(s.counters["last_nweno"] / ncell)
            if time.time() - t_log > 30.0:
                t_log = time.time()
                log(f"[evaluate] {name} {mode}: t = {s.time[0]:.4f}, {len(fr)} steps")
        s.sync()
        el = time.time() - t0
        c = s.counters
        states[mode] = s.state
        res[mode] = {"seconds": el, "rhs_evals": c["rhs_evals"], "steps": s.time[2] - 1,
                     "mzone_updates_per_s": ncell * c["rhs_evals"] / el / 1e6,
                     "weno_fraction": float(np.mean(fr)) if fr else 1.0, "t_final": s.time[0]}
        if mode == "learned":
            res[mode].update(s.indicator_counters)
        s.close()
        log(f"[evaluate] {name} {mode}: {res[mode]}")
    if "full" in states:
        for mode in ("kxrcf", "learned"):
            if mode in states:
                res[mode]["l1_vs_full_weno"] = float(np.abs(states[mode] - states["full"]).mean())
        if "kxrcf" in states and "learned" in states:
            res["l1_ratio_learned_vs_kxrcf"] = (res["learned"]["l1_vs_full_weno"] /
                                                res["kxrcf"]["l1_vs_full_weno"])
    elif "kxrcf" in states and "learned" in states:
        res["learned"]["l1_vs_kxrcf"] = float(np.abs(states["learned"] - states["kxrcf"]).mean())
    # wall time to the same end time (redone steps count against the learned run)
    for ref in ("kxrcf", "full"):
        if ref in res and "learned" in res:
            res[f"speedup_learned_vs_{ref}"] = res[ref]["seconds"] / res["learned"]["seconds"]
    return res


def load(name):
    """bfloat16 weights (uint16) of a trained recipe."""
    return np.load(os.path.join(WEIGHTS, f"{name}.npy"), allow_pickle=False)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("recipe", choices=sorted(RECIPES))
    ap.add_argument("--out", default=WEIGHTS, help="directory of the trained weights")
    ap.add_argument("--eval-out", default=".", help="directory of the evaluation's JSON result")
    ap.add_argument("--evaluate", action="store_true", help="also run the recipe's evaluation problem")
    ap.add_argument("--evaluate-only", action="store_true")
    ap.add_argument("--n", type=int, default=0, help="evaluation resolution per axis (default: the recipe's)")
    ap.add_argument("--modes", default="full,kxrcf,learned", help="runs of the evaluation, comma-separated")
    a = ap.parse_args()
    W = None
    if not a.evaluate_only:
        W, _ = train(a.recipe, a.out)
    if a.evaluate or a.evaluate_only:
        dim = EVALUATIONS[a.recipe]["dim"]
        modes = tuple(a.modes.split(","))
        res = evaluate(a.recipe, W, n=(a.n,) * dim if a.n else None, modes=modes)
        tag = (f"_n{a.n}" if a.n else "") + ("" if "full" in modes else "_sparse")
        with open(os.path.join(a.eval_out, f"{a.recipe}_evaluation{tag}.json"), "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
