"""Batches of independent 2D problems (kfvm_create_batch): each problem of a
batch evolves exactly as the same problem run alone, with fixed and with
adaptive step sizes, and an ensemble of uneven sizes reproduces its members."""
import numpy as np
import pytest

from helpers import config, init_text

pytestmark = pytest.mark.gpu


def _single(c, seed):
    import kfvm_amd
    return kfvm_amd.Solver(init_text=init_text(dict(c, seed=seed)))


def test_fixed_dt_batch_matches_single_runs():
    import kfvm_amd
    c = config(dim=2, n=(40, 36, 1), problem="quadrant2d", user=(), fixeddt=0.002)
    seeds = [3, 7, 11]
    b = kfvm_amd.Solver(init_text=init_text(c), seeds=seeds)
    assert b.shape == (5, 3, 36, 40)
    singles = [_single(c, s) for s in seeds]
    for _ in range(5):
        assert b.take_step() == 0
        for s in singles:
            assert s.take_step() == 0
    U = b.state
    for k, s in enumerate(singles):
        d = np.abs(U[:, k:k + 1] - s.state).max()
        print(f"[batch] fixed dt, problem {k}: max diff {d:.3e}", flush=True)
        assert d == 0.0
    t, dt = b.slice_times
    assert np.all(t == singles[0].time[0])


def test_adaptive_batch_matches_single_runs():
    import kfvm_amd
    c = config(dim=2, n=(48, 48, 1), problem="quadrant2d", user=(), final=0.04)
    seeds = [1, 5, 9, 13]
    b = kfvm_amd.Solver(init_text=init_text(c), seeds=seeds)
    b.solve()
    t, _ = b.slice_times
    U = b.state
    for k, sd in enumerate(seeds):
        s = _single(c, sd)
        s.solve()
        cnt = s.counters
        d = np.abs(U[:, k:k + 1] - s.state).max()
        print(f"[batch] adaptive, problem {k}: t {t[k]:.6f} vs {s.time[0]:.6f}, max diff {d:.3e}, "
              f"rejections {cnt['rejects_tol']}+{cnt['rejects_unphys']}", flush=True)
        assert t[k] == s.time[0]
        if cnt["rejects_tol"] + cnt["rejects_unphys"] == 0:
            assert d == 0.0
        else:
            assert d < 1e-2 * np.abs(s.state).max()


def test_ensemble_of_uneven_sizes_reproduces_its_members():
    from kfvm_amd.ensemble import Ensemble, quadrant_text
    import kfvm_amd
    problems = [(32, 2), (48, 3), (32, 4), (48, 5), (64, 6)]
    e = Ensemble(problems, lambda n: quadrant_text(n, 0.02))
    el, zones = e.run()
    assert zones > 0
    for k, (n, sd) in enumerate(problems):
        s = kfvm_amd.Solver(init_text=quadrant_text(n, 0.02).replace("problem:bc outflow", "problem:bc outflow\nuser:seed %d" % sd))
        s.solve()
        d = np.abs(e.state(k) - s.state).max()
        print(f"[batch] ensemble member {k} ({n}^2): max diff {d:.3e}", flush=True)
        assert d < 1e-2 * np.abs(s.state).max()
    e.close()
