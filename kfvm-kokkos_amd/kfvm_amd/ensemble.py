"""Batched ensembles of independent 2D problems of uneven sizes.

Problems of one grid size form one batch (kfvm_create_batch): every kernel of
an RK stage covers all of the batch's problems in a single launch, and each
problem keeps its own step size and controller.  The batches of different
sizes run on their own streams, driven from host threads, so they overlap on
the GPU.

    python -m kfvm_amd.ensemble [--problems 512] [--t-end 0.25] [--sizes 64 128 256 512]

runs the four-quadrant Riemann ensemble (problem k: seed k, grid size drawn
uniformly from --sizes by a seeded generator; quadrant states rho, p in
U[0.1, 2], u, v in U[-1, 1]; split at (0.5, 0.5); outflow; HLLC) and prints
one JSON line with the aggregate zone-update rate.
"""
import argparse
import json
import threading
import time

import numpy as np

import kfvm_amd


def quadrant_text(n, t_end, nq=2, wenothresh=1.0):
    return "\n".join([
        f"mesh:resolution {n} {n}", "mesh:extent 0.0 1.0 0.0 1.0", "stencil:lfac 5",
        f"time:final {t_end!r}", "time:atol 1.e-3", "time:rtol 1.e-3", "time:maxsteps 100000",
        "eos:gamma 1.4", f"eos:wenothresh {wenothresh!r}", "problem:name quadrant2d", "problem:dim 2",
        f"problem:nquad {nq}", "problem:equation hydro", "problem:riemann hllc", "problem:rk ssp4_3_2",
        "problem:bc outflow"]) + "\n"


class Ensemble:
    """problems: list of (grid size n, seed); text(n) gives the init text of a
    problem of size n.  tuning: {size: flux brick shape} (default: the table of
    kfvm_amd.autotune, when one was written; {} leaves every batch at the
    default 16x16 bricks)."""

    def __init__(self, problems, text, tuning=None):
        from kfvm_amd.autotune import load_table, shape_for
        table = load_table() if tuning is None else tuning
        self.groups = {}
        for k, (n, seed) in enumerate(problems):
            self.groups.setdefault(int(n), []).append((k, int(seed)))
        self.shapes = {n: shape_for(n, table) for n in self.groups}
        self.batches = {n: kfvm_amd.Solver(init_text=text(n) + f"tune:fluxbrick {self.shapes[n]}\n",
                                           seeds=[sd for _, sd in g])
                        for n, g in sorted(self.groups.items())}

    def run(self):
        """Step every batch to its final time, batches concurrently."""
        errors = []

        def drive(s):
            try:
                s.solve()
            except Exception as e:  # noqa: BLE001 - reported below
                errors.append(e)

        t0 = time.perf_counter()
        th = [threading.Thread(target=drive, args=(s,)) for s in self.batches.values()]
        for t in th:
            t.start()
        for t in th:
            t.join()
        for s in self.batches.values():
            s.sync()
        el = time.perf_counter() - t0
        if errors:
            raise errors[0]
        zones = sum(int(np.prod(s.shape[1:])) * s.counters["rhs_evals"] for s in self.batches.values())
        return el, zones

    def state(self, k):
        """Final state [NV, 1, n, n] of problem k."""
        for n, g in self.groups.items():
            for pos, (kk, _) in enumerate(g):
                if kk == k:
                    return self.batches[n].state[:, pos:pos + 1]
        raise KeyError(k)

    def close(self):
        for s in self.batches.values():
            s.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--problems", type=int, default=512)
    ap.add_argument("--t-end", type=float, default=0.25)
    ap.add_argument("--sizes", type=int, nargs="+", default=[64, 128, 256, 512])
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args()
    rng = np.random.default_rng(a.seed)
    sizes = rng.choice(a.sizes, a.problems)
    e = Ensemble([(int(n), k) for k, n in enumerate(sizes)], lambda n: quadrant_text(n, a.t_end))
    el, zones = e.run()
    per = {int(n): int((sizes == n).sum()) for n in a.sizes}
    steps = {n: s.time[2] - 1 for n, s in e.batches.items()}
    print(json.dumps({"problems": a.problems, "cells": int(sum(int(n) ** 2 for n in sizes)), "per_size": per,
                      "t_end": a.t_end, "seconds": el, "zone_updates": zones,
                      "mzone_updates_per_s": zones / el / 1e6, "steps_per_batch": steps,
                      "flux_brick": {n: e.shapes[n] for n in sorted(e.shapes)}}))
    e.close()


if __name__ == "__main__":
    main()
