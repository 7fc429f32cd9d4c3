"""Batches of independent 2D problems (kfvm_create_batch): each problem of a
batch evolves bit for bit as the same problem run alone, with fixed and with
adaptive step sizes (rejected attempts of one problem re-evaluate the others
with a zero step, which must leave their state, limiter thetas and WENO flags
untouched, and an unphysical attempt forces full WENO on its own problem
only), an ensemble of uneven sizes reproduces its members, and one member
agrees with the CPU oracle run of the same problem."""
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


@pytest.mark.parametrize("kind", ["tolerance", "unphysical"])
def test_adaptive_batch_matches_single_runs(kind):
    """Adaptive steps: tight tolerances make members reject attempts for
    accuracy (some to the rejection limit, where a member stops with its last
    attempt's state); a first step about 5 times the stable one makes members
    reject attempts as unphysical (the retry evaluates that member with full
    WENO only)."""
    import kfvm_amd
    if kind == "tolerance":
        c = config(dim=2, n=(48, 48, 1), problem="quadrant2d", user=(), final=0.04, atol=1e-6, rtol=1e-6)
    else:
        c = config(dim=2, n=(48, 48, 1), problem="quadrant2d", user=(), final=0.04, cfl=50.0)
    seeds = [1, 5, 9, 13]
    b = kfvm_amd.Solver(init_text=init_text(c), seeds=seeds)
    first_dt = 0.04 if kind == "unphysical" else None  # about 5 times the stable step
    if first_dt:
        b.set_time(0.0, first_dt)
    b.solve()
    t, _ = b.slice_times
    U = b.state
    rej = 0
    for k, sd in enumerate(seeds):
        s = _single(c, sd)
        if first_dt:
            s.set_time(0.0, first_dt)
        s.solve()
        cnt = s.counters
        d = np.abs(U[:, k:k + 1] - s.state).max()
        print(f"[batch] adaptive ({kind}), problem {k}: t {t[k]:.6f} vs {s.time[0]:.6f}, max diff {d:.3e}, "
              f"rejections {cnt['rejects_tol']}+{cnt['rejects_unphys']}", flush=True)
        assert t[k] == s.time[0]
        assert d == 0.0
        assert np.array_equal(b.flags[:, k:k + 1], s.flags)
        rej += cnt["rejects_tol"] if kind == "tolerance" else cnt["rejects_unphys"]
    # the batch's retries are exercised: some member rejected an attempt
    assert rej > 0


def test_ensemble_of_uneven_sizes_reproduces_its_members():
    from kfvm_amd.ensemble import Ensemble, quadrant_text
    import kfvm_amd
    problems = [(32, 2), (48, 3), (32, 4), (48, 5), (64, 6)]
    e = Ensemble(problems, lambda n: quadrant_text(n, 0.02), tuning={})  # the single runs' 16x16 bricks
    el, zones = e.run()
    assert zones > 0
    for k, (n, sd) in enumerate(problems):
        s = kfvm_amd.Solver(init_text=quadrant_text(n, 0.02).replace("problem:bc outflow", "problem:bc outflow\nuser:seed %d" % sd))
        s.solve()
        d = np.abs(e.state(k) - s.state).max()
        print(f"[batch] ensemble member {k} ({n}^2): max diff {d:.3e}", flush=True)
        assert d == 0.0
    e.close()


def test_batch_member_matches_the_oracle():
    """Member k of a fixed-dt batch against the oracle run of problem k (same
    seed), each from its own initial condition: 3 steps, per-cell relative
    difference <= 1e-12."""
    import kfvm_amd
    from helpers import cell_rel_diff, oracle_for
    c = config(dim=2, n=(40, 36, 1), problem="quadrant2d", user=(), fixeddt=0.002)
    seeds = [3, 7, 11]
    b = kfvm_amd.Solver(init_text=init_text(c), seeds=seeds)
    k = 1
    text = init_text(dict(c, seed=seeds[k]))
    p = kfvm_amd.parse_init(text)
    f, dw = b.weights
    o = oracle_for(c, p, f, dw)
    o.set_ic(c["problem"], p["user"], p["seed"])
    assert np.array_equal(o.state, b.state[:, k:k + 1])
    fl = np.zeros((2,) + o.state.shape[1:])
    fl[1] = 1.0
    o.flags = fl
    o.prime()
    for _ in range(3):
        assert b.take_step() == 0 and o.take_step() == 0
    r = cell_rel_diff(b.state[:, k:k + 1], o.state).max()
    print(f"[batch] member {k} vs oracle after 3 steps: max cell rel diff {r:.3e}", flush=True)
    assert r <= 1e-12


@pytest.mark.parametrize("kind", ["fixed", "tolerance"])
def test_graph_replayed_steps_equal_eager_steps(kind, monkeypatch, capfd):
    """Step attempts of a batch captured once as HIP graphs and replayed
    (KFVM_GRAPH=1, the default) against the same launches issued one by one
    (KFVM_GRAPH=0): after every step, with the host's own reads and
    synchronisations between the replays, states, flags and times are bit for
    bit equal; the tolerance case replays the re-priming attempts of
    rejected members too."""
    import kfvm_amd
    if kind == "fixed":
        c = config(dim=2, n=(40, 36, 1), problem="quadrant2d", user=(), fixeddt=0.002)
    else:
        c = config(dim=2, n=(48, 48, 1), problem="quadrant2d", user=(), final=0.02, atol=1e-6, rtol=1e-6)
    seeds = [2, 4, 6, 8]
    monkeypatch.setenv("KFVM_GRAPH", "0")
    eager = kfvm_amd.Solver(init_text=init_text(c), seeds=seeds)
    monkeypatch.setenv("KFVM_GRAPH", "1")
    graph = kfvm_amd.Solver(init_text=init_text(c), seeds=seeds)
    steps = 0
    for _ in range(12):
        se, sg = eager.take_step(), graph.take_step()
        assert se == sg
        steps += 1
        assert np.array_equal(graph.state, eager.state)
        assert np.array_equal(graph.flags, eager.flags)
        te, dte = eager.slice_times
        tg, dtg = graph.slice_times
        assert np.array_equal(te, tg) and np.array_equal(dte, dtg)
        if graph.time[0] >= c["final"]:
            break
    ce, cg = eager.counters, graph.counters
    print(f"[batch] graph replay ({kind}): {steps} steps equal, RHS evaluations {cg['rhs_evals']} / "
          f"{ce['rhs_evals']}, rejections {cg['rejects_tol']}", flush=True)
    assert ce == cg
    if kind == "tolerance":
        assert cg["rejects_tol"] > 0
    # the graphs were captured (a failed capture warns and runs eagerly)
    assert "graph capture" not in capfd.readouterr().err


@pytest.mark.parametrize("case", ["euler_batch_nq2", "mhd_nq3", "linadv"])
def test_flux_brick_shapes_give_identical_states(case):
    """The 2D flux kernel's brick shapes (tune:fluxbrick 0 = 16x16, 1 = 32x8,
    2 = 64x4) tile the same faces: with a fixed step, three steps give the same
    states bit for bit, on grids that are no multiple of the bricks."""
    import kfvm_amd
    from kfvm_amd.ensemble import quadrant_text
    if case == "euler_batch_nq2":
        base, seeds = quadrant_text(48, 1.0) + "time:fixeddt 0.002\n", [1, 2, 3]
    else:
        c = (config(dim=2, n=(40, 36, 1), nq=3, equation="mhd_glm", riemann="kepes", problem="orszag_tang",
                    user=(), bc="periodic", gamma=5.0 / 3.0, fixeddt=0.001) if case == "mhd_nq3" else
             config(dim=2, n=(44, 30, 1), equation="linadv", riemann="exact", problem="linadv_sine",
                    user=(1.0, 1.0, 0.0), bc="periodic", advvel=(1.0, 1.0, 0.0), fixeddt=0.004))
        base, seeds = init_text(c), None
    states = []
    for shape in (0, 1, 2):
        s = kfvm_amd.Solver(init_text=base + f"tune:fluxbrick {shape}\n", seeds=seeds)
        for _ in range(3):
            assert s.take_step() == 0
        states.append(s.state)
        s.close()
    for shape in (1, 2):
        d = np.abs(states[shape] - states[0]).max()
        print(f"[batch] {case}: brick shape {shape} vs 16x16: max diff {d:.3e}", flush=True)
        assert d == 0.0
