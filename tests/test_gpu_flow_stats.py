"""Flow statistics on the GPU (kfvm_flow_stats, Solver::evalFlowStats) against
the restatement helpers.flow_stats_ref (pinned by tests/test_flow_stats.py) on
the same states, and the stats file kfvm_solve writes every io:statsfreq
steps (rows of <basename>_stats.h5, dataset stats_data, read back with
h5dump).  Sums differ from the restatement's in order only: 1e-12 relative."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from helpers import config, flow_stats_ref, init_text

pytestmark = pytest.mark.gpu

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)

CASES = {
    "euler3d_outflow": config(n=(12, 10, 9)),
    "euler3d_reflect_noslip": config(n=(10, 8, 8), bc=("reflecting", "noslip", "reflecting", "outflow",
                                                        "noslip", "reflecting")),
    "euler2d_periodic": config(dim=2, n=(24, 20, 1), problem="shear2d", user=(), bc="periodic"),
    "mhd3d_periodic": config(n=(10, 10, 10), equation="mhd_glm", riemann="llf", problem="mhd_vortex3d", user=(),
                             bc="periodic", gamma=5.0 / 3.0),
    "mhd2d_reflect": config(dim=2, n=(20, 16, 1), equation="mhd_glm", riemann="hll", problem="mhd_blast",
                            user=(0.1, 0.70710678118654752, 0.70710678118654752), bc="reflecting",
                            gamma=5.0 / 3.0),
}


def _random_state(rng, nv, shape):
    U = rng.uniform(-1.0, 1.0, (nv,) + shape)
    U[0] = rng.uniform(0.5, 2.0, shape)
    U[4] = rng.uniform(8.0, 12.0, shape)
    return U


@pytest.mark.parametrize("name", sorted(CASES))
def test_flow_stats_match_the_restatement(name):
    import kfvm_amd
    c = CASES[name]
    s = kfvm_amd.Solver(init_text=init_text(c))
    p = kfvm_amd.parse_init(init_text(c))
    dim = p["dim"]
    n = p["n"]
    spacing = tuple((p["hi"][a] - p["lo"][a]) / n[a] for a in range(3))
    volume = float(np.prod([p["hi"][a] - p["lo"][a] for a in range(dim)]))
    rng = np.random.default_rng(11)
    for U in (s.state, _random_state(rng, s.state.shape[0], s.state.shape[1:])):
        s.state = U
        got = s.flow_stats
        ref = flow_stats_ref(U, dim, spacing, volume, p["bc"], c["equation"])
        rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
        print(f"[stats] {name}: {got} vs {ref}, rel {rel.max():.2e}", flush=True)
        assert got.shape == ref.shape and rel.max() <= 1e-12
    s.close()


def _h5(path, name):
    out = subprocess.run([H5DUMP, "-d", name, "-y", "-w", "0", "-m", "%.17g", path], check=True,
                         capture_output=True, text=True).stdout
    dims = re.search(r"DATASPACE\s+SIMPLE \{ \( ([0-9, ]+) \)", out)
    shape = tuple(int(x) for x in dims.group(1).split(","))
    body = out[out.index("DATA {") + 6:]
    vals = np.array([float(x) for x in re.findall(r"-?[0-9.]+(?:e[-+]?[0-9]+)?", body.split("}")[0])])
    return shape, vals.reshape(shape)


@pytest.mark.skipif(H5DUMP is None, reason="needs h5dump")
def test_solve_writes_a_stats_row_every_statsfreq_steps(tmp_path):
    """Solver::Solve's sequence (Solver.cpp:88-127): a row before the first
    step, one after every step whose number is a multiple of io:statsfreq,
    one after the last step (both when the last step is such a multiple)."""
    import kfvm_amd
    c = config(n=(12, 12, 12), fixeddt=0.004, final=0.019, maxsteps=40)
    text = init_text(c) + f"io:datadir {tmp_path}\nio:basename run\nio:statsfreq 2\n"
    s = kfvm_amd.Solver(init_text=text)
    first = s.flow_stats
    s.solve()
    last = s.flow_stats
    t, _, step = s.time
    s.close()
    path = tmp_path / "R2_NX12_NY12_NZ12" / "run_stats.h5"
    shape, rows = _h5(str(path), "stats_data")
    assert shape == (40, 3)
    # 4 steps of 0.004 and one of 0.003: rows at t = 0, after steps 2 and 4,
    # and after the last (5)
    times = rows[:, 0]
    nrow = 4
    print(f"[stats] rows: {rows[:nrow + 1]}", flush=True)
    assert np.all(rows[nrow:] == 0.0)
    # time:final is read with std::stof semantics (the reference's parser): the
    # last row's time is the solver's final time, 0.019 rounded to float
    np.testing.assert_allclose(times[:nrow], [0.0, 0.008, 0.016, t], rtol=1e-12, atol=1e-15)
    assert np.array_equal(rows[0, 1:], first) and np.array_equal(rows[nrow - 1, 1:], last)
    assert t == float(np.float32(0.019))
