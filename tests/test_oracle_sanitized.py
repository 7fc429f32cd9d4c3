"""The CPU oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
(oracle/Makefile target `sanitize`, host code only) runs every checked kind of
configuration -- Euler 2D/3D with sparse and full WENO, reflecting and
periodic walls, NQ = 3, linear advection, GLM-MHD with HLL / KEPES -- for a few
adaptive steps, and every other entry point once, without a report.

A child process preloads the ASan runtime and loads the sanitized library
through KFVM_ORACLE_LIB; the weights are built here and handed over as .npy
files.  Catches out-of-bounds stencil or halo indexing, use of freed or
uninitialised buffers in the restated kernels, and signed overflow / invalid
shifts in the index arithmetic, none of which the value checks would see.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import config, init_text, mhd_tables

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

CASES = {
    "euler3d_sparse": config(n=(8, 8, 8)),
    "euler3d_full_reflect": config(n=(8, 6, 5), wenothresh=-1.0, bc="reflecting", riemann="hll"),
    "euler3d_periodic_nq3": config(n=(7, 6, 5), nq=3, bc="periodic", rk="fourstarp"),
    "quadrant2d_noslip": config(dim=2, n=(12, 10, 1), problem="quadrant2d", seed=3, user=(), bc="noslip",
                                riemann="llf", rk="threestarp"),
    "linadv2d_nq3": config(dim=2, n=(12, 12, 1), nq=3, equation="linadv", riemann="exact",
                           problem="linadv_sine", user=(1.0, 1.0, 0.0), bc="periodic", wenothresh=-1.0,
                           advvel=(1.0, 1.0, 0.0), rk="ssp10_4_3"),
    "mhd3d_kepes": config(n=(6, 6, 6), equation="mhd_glm", riemann="kepes", problem="mhd_vortex3d", user=(),
                          bc="periodic", gamma=5.0 / 3.0),
    "orszag_tang2d_hll_nq3": config(dim=2, n=(10, 10, 1), nq=3, equation="mhd_glm", riemann="hll",
                                    problem="orszag_tang", user=(), bc="periodic", gamma=5.0 / 3.0),
}

CHILD = r'''
import json, os, sys
import numpy as np
sys.path.insert(0, os.path.join(sys.argv[1], "oracle"))
from orc import Oracle
d = sys.argv[2]
spec = json.load(open(os.path.join(d, "cases.json")))
for name, kw in spec.items():
    arr = {k: np.load(os.path.join(d, f"{name}.{k}.npy")) for k in ("face_w", "deriv_w", "cell_w", "diff_m")
           if os.path.exists(os.path.join(d, f"{name}.{k}.npy"))}
    problem, user, seed = kw.pop("problem"), kw.pop("user"), kw.pop("seed")
    o = Oracle(threads=2, **kw, **arr)
    o.set_ic(problem, user, seed)
    o.prime()
    for _ in range(3):
        o.take_step()
    o.rhs, o.flags, o.state, o.time
    print(name, "ok", flush=True)
'''


def _oracle_kwargs(c, p):
    kw = dict(dim=p["dim"], n=list(p["n"][: p["dim"]]), nq=p["nq"], eq=c["equation"], rs=c["riemann"],
              rk=c["rk"], bc=p["bc"], lo=list(p["lo"]), hi=list(p["hi"]), gamma=p["gamma"],
              weno_thresh=p["wenothresh"], adv_vel=p["advvel"], cfl=p["cfl"] if p["cfl"] > 0 else None,
              final_time=p["final"], min_dt=p["mindt"], atol=p["atol"], rtol=p["rtol"],
              rejection_thresh=p["rejectionthresh"], rejection_limit=p["rejectionlimit"],
              max_steps=p["maxsteps"], fixed_dt=p["fixeddt"])
    if kw["adv_vel"] is not None:
        kw["adv_vel"] = [float(v) for v in kw["adv_vel"]]
    kw.update(problem=p["problem"], user=[float(u) for u in p["user"]], seed=int(p["seed"]))
    return kw


def test_oracle_under_address_and_ub_sanitizers(tmp_path):
    import kfvm_amd
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    spec = {}
    for name, c in CASES.items():
        p = kfvm_amd.parse_init(init_text(c))
        n = p["n"]
        dx = [(p["hi"][a] - p["lo"][a]) / n[a] for a in range(3)]
        fw, dw = kfvm_amd.build_weights(p["dim"], p["nq"], p["lfac"], dx[0], dx[1], dx[2] if p["dim"] == 3 else 1.0)
        np.save(tmp_path / f"{name}.face_w.npy", fw)
        np.save(tmp_path / f"{name}.deriv_w.npy", dw)
        if c["equation"] == "mhd_glm":
            cw, dm = mhd_tables(p)
            np.save(tmp_path / f"{name}.cell_w.npy", cw)
            np.save(tmp_path / f"{name}.diff_m.npy", dm)
        spec[name] = _oracle_kwargs(c, p)
    (tmp_path / "cases.json").write_text(json.dumps(spec))
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, KFVM_ORACLE_LIB=os.path.join(ORACLE, "build", "san", "liborc.so"),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:exitcode=23", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT, str(tmp_path)], env=env, capture_output=True,
                       text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert all(f"{name} ok" in r.stdout for name in CASES), out[-4000:]
