"""Tile autotuner of the batched 2D path.

Times the flux kernel's brick shapes (init key tune:fluxbrick: 0 = 16x16,
1 = 32x8, 2 = 64x4 cells, 256 per workgroup in each) on batches of every grid
size and records the fastest shape per size:

    python -m kfvm_amd.autotune [--sizes 64 128 256 512] [--problems 64] [--steps 4] [--out FILE]

For each size and shape: a batch of `problems` four-quadrant problems (seeds
0, 1, ...), two untimed steps, then `steps` timed steps with the per-kernel
HIP-event timing; the flux kernel's milliseconds per launch decide.  With a
fixed step the shapes give the same states bit for bit (only the brick tiling
of the faces changes; with the adaptive controller the error norm's block
partial sums regroup, a rounding-level difference), so the choice is a speed
choice.  The table is written as JSON
(default: kfvm_amd/tuning/flux_brick_2d.json), which ensemble.Ensemble
applies to each size class.
"""
import argparse
import json
import os
import sys

import numpy as np

import kfvm_amd
from kfvm_amd.ensemble import quadrant_text

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning", "flux_brick_2d.json")
SHAPES = {0: "16x16", 1: "32x8", 2: "64x4"}


def load_table(path=TABLE):
    """{grid size: brick shape} of a tuning run, {} when there is none."""
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return {int(k): int(v) for k, v in json.load(f)["best"].items()}


def shape_for(n, table):
    """The tuned shape of size n, or that of the nearest tuned size (0 without a table)."""
    if not table:
        return 0
    return table[min(table, key=lambda m: abs(np.log2(m) - np.log2(n)))]


def time_shape(n, shape, problems, steps, t_end=0.25):
    # a fixed step (the same stage sequence as the adaptive run): the states
    # of the shapes are then comparable bit for bit
    text = quadrant_text(n, t_end) + f"tune:fluxbrick {shape}\ntime:fixeddt 1e-4\n"
    s = kfvm_amd.Solver(init_text=text, seeds=list(range(problems)))
    for _ in range(2):
        s.take_step()
    s.sync()
    s.profile(True)
    s.profile_reset()
    for _ in range(steps):
        s.take_step()
    s.sync()
    ms, cnt = s.profile_get("flux_rk")
    stage = sum(s.profile_get(k)[0] for k in ("cell_props", "linear", "select", "weno", "weno_post", "pp",
                                            "flux_rk", "stage_pre"))
    s.profile(False)
    state = s.state
    s.close()
    return ms / max(cnt, 1), stage / max(cnt, 1), state


def tune(sizes, problems, steps, log=print):
    res = {}
    for n in sizes:
        rows = {}
        ref = None
        for shape in SHAPES:
            flux, stage, state = time_shape(n, shape, problems, steps)
            same = ref is None or bool(np.array_equal(state, ref))
            ref = state if ref is None else ref
            rows[shape] = {"flux_ms": flux, "stage_ms": stage, "same_states": same}
            log(f"[autotune] {n}^2 x {problems}: brick {SHAPES[shape]}: flux {flux:.3f} ms, "
                f"stage {stage:.3f} ms, states equal to 16x16: {same}")
        best = min(rows, key=lambda k: rows[k]["flux_ms"])
        res[n] = {"best": best, "shapes": rows}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", type=int, nargs="+", default=[64, 128, 256, 512])
    ap.add_argument("--problems", type=int, default=64)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--out", default=TABLE)
    a = ap.parse_args()
    res = tune(a.sizes, a.problems, a.steps, log=lambda m: print(m, file=sys.stderr, flush=True))
    out = {"problems": a.problems, "steps": a.steps, "shapes": SHAPES,
           "best": {str(n): r["best"] for n, r in res.items()},
           "flux_ms": {str(n): {SHAPES[k]: v["flux_ms"] for k, v in r["shapes"].items()} for n, r in res.items()},
           "stage_ms": {str(n): {SHAPES[k]: v["stage_ms"] for k, v in r["shapes"].items()} for n, r in res.items()},
           "same_states": all(v["same_states"] for r in res.values() for v in r["shapes"].values())}
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
