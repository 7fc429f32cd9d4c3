"""ctypes binding of the CPU oracle (oracle/build/liborc.so).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the HIP library
and by bench.py's cpu_baseline leg.  See oracle.h for what it restates.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# KFVM_ORACLE_LIB: another build of the same source (the sanitizer build of
# tests/test_oracle_sanitized.py)
LIB_PATH = os.environ.get("KFVM_ORACLE_LIB") or os.path.join(HERE, "build", "liborc.so")

EQ = {"hydro": 0, "linadv": 1, "mhd_glm": 2}
RS = {"llf": 0, "hll": 1, "hllc": 2, "exact": 3, "kepes": 4}
RK = {"threestarp": 0, "fourstarp": 1, "ssp4_3_2": 2, "ssp10_4_3": 3}
BC = {"periodic": 0, "outflow": 1, "reflecting": 2, "noslip": 3}


class OrcConfig(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("nq", ctypes.c_int), ("eq", ctypes.c_int), ("rs", ctypes.c_int),
        ("rk", ctypes.c_int), ("bc", ctypes.c_int * 6), ("n", ctypes.c_int64 * 3),
        ("lo", ctypes.c_double * 3), ("hi", ctypes.c_double * 3),
        ("gamma", ctypes.c_double), ("wenoThresh", ctypes.c_double),
        ("haveAdvVel", ctypes.c_int), ("advVel", ctypes.c_double * 3),
        ("cfl", ctypes.c_double), ("finalTime", ctypes.c_double), ("minDt", ctypes.c_double),
        ("atol", ctypes.c_double), ("rtol", ctypes.c_double), ("rejectionThresh", ctypes.c_double),
        ("fixedDt", ctypes.c_double), ("rejectionLimit", ctypes.c_int), ("maxSteps", ctypes.c_int),
        ("nThreads", ctypes.c_int),
    ]


_lib = None


def build():
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P, D, I64 = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int64
        L.orc_create.argtypes = [ctypes.POINTER(OrcConfig), D, D]
        L.orc_create.restype = P
        L.orc_destroy.argtypes = [P]
        L.orc_destroy.restype = None
        for n in ("orc_set_state", "orc_get_state", "orc_get_rhs", "orc_get_flags", "orc_set_flags"):
            getattr(L, n).argtypes = [P, D]
            getattr(L, n).restype = None
        L.orc_set_ic.argtypes = [P, ctypes.c_char_p, D, ctypes.c_uint64]
        L.orc_set_ic.restype = ctypes.c_int
        L.orc_eval_rhs.argtypes = [P, ctypes.c_int]
        L.orc_eval_rhs.restype = ctypes.c_double
        L.orc_prime.argtypes = [P]
        L.orc_prime.restype = None
        L.orc_take_step.argtypes = [P]
        L.orc_take_step.restype = ctypes.c_int
        L.orc_get_time.argtypes = [P, D, D, ctypes.POINTER(I64), D]
        L.orc_get_time.restype = None
        L.orc_set_time.argtypes = [P, ctypes.c_double, ctypes.c_double]
        L.orc_set_time.restype = None
        L.orc_rhs_evals.argtypes = [P]
        L.orc_rhs_evals.restype = I64
        L.orc_force_full_weno.argtypes = [P]
        L.orc_force_full_weno.restype = None
        L.orc_num_weno.argtypes = [P]
        L.orc_num_weno.restype = I64
        L.orc_last_step.argtypes = [P]
        L.orc_last_step.restype = ctypes.c_int
        L.orc_set_mhd_tables.argtypes = [P, D, D]
        L.orc_set_mhd_tables.restype = None
        L.orc_face_states.argtypes = [P, ctypes.c_int, D]
        L.orc_face_states.restype = None
        L.orc_weno_omegas.argtypes = [P, I64, I64, I64, D]
        L.orc_weno_omegas.restype = None
        L.orc_error_norm.argtypes = [ctypes.c_int, ctypes.c_int, I64, D, D, ctypes.c_double, ctypes.c_double, D, D]
        L.orc_error_norm.restype = None
        L.orc_riemann_point.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, D, D, D]
        L.orc_riemann_point.restype = ctypes.c_double
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def riemann_point(eq, rs, gamma, ch, direction, uL, uR):
    """One face-point Riemann solve: (flux, max wave speed)."""
    L = lib()
    a = np.ascontiguousarray(uL, dtype=np.float64)
    b = np.ascontiguousarray(uR, dtype=np.float64)
    F = np.zeros_like(a)
    s = L.orc_riemann_point(EQ[eq], RS[rs], gamma, ch, direction, _dp(a), _dp(b), _dp(F))
    return F, s


def error_norm(eq, U, Uhat, atol, rtol):
    """RKFSAL_ErrEst_K over cells: (sum of squared scaled differences, min
    positivity).  U, Uhat: [ncell, nv]."""
    L = lib()
    a = np.ascontiguousarray(U, dtype=np.float64)
    b = np.ascontiguousarray(Uhat, dtype=np.float64)
    assert a.shape == b.shape and a.ndim == 2
    e, p = ctypes.c_double(), ctypes.c_double()
    L.orc_error_norm(a.shape[1], EQ[eq], a.shape[0], _dp(a), _dp(b), atol, rtol, ctypes.byref(e), ctypes.byref(p))
    return e.value, p.value


class Oracle:
    """CPU restatement of KFVM::Solver for the checked subset of the path."""

    def __init__(self, *, dim, n, face_w, deriv_w, nq=2, eq="hydro", rs="hllc", rk="ssp4_3_2",
                 bc="outflow", lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), gamma=1.4,
                 weno_thresh=1.0, adv_vel=None, cfl=None, final_time=1.0, min_dt=1e-10,
                 atol=1e-4, rtol=1e-4, rejection_thresh=0.8, rejection_limit=4, max_steps=1000,
                 fixed_dt=-1.0, threads=0, cell_w=None, diff_m=None):
        self.L = lib()
        c = OrcConfig()
        c.dim, c.nq, c.eq, c.rs, c.rk = dim, nq, EQ[eq], RS[rs], RK[rk]
        bcs = [bc] * 6 if isinstance(bc, str) else list(bc)
        for f in range(6):
            c.bc[f] = BC[bcs[f]]
        nn = list(n) + [1] * (3 - len(n))
        for d in range(3):
            c.n[d] = nn[d]
            c.lo[d], c.hi[d] = lo[d], hi[d]
        c.gamma, c.wenoThresh = gamma, weno_thresh
        c.haveAdvVel = 1 if adv_vel is not None else 0
        if adv_vel is not None:
            for d in range(3):
                c.advVel[d] = adv_vel[d] if d < len(adv_vel) else 0.0
        c.cfl = cfl if cfl is not None else (0.5 if dim == 3 else 0.75)
        c.finalTime, c.minDt, c.atol, c.rtol = final_time, min_dt, atol, rtol
        c.rejectionThresh, c.fixedDt = rejection_thresh, fixed_dt
        c.rejectionLimit, c.maxSteps, c.nThreads = rejection_limit, max_steps, threads
        self.cfg = c
        self.nv = {"hydro": 5, "linadv": 1, "mhd_glm": 9}[eq]
        self.shape = (self.nv, nn[2] if dim == 3 else 1, nn[1], nn[0])
        self._fw = np.ascontiguousarray(face_w, dtype=np.float64)
        self._dw = np.ascontiguousarray(deriv_w, dtype=np.float64)
        self.h = self.L.orc_create(ctypes.byref(c), _dp(self._fw), _dp(self._dw))
        if not self.h:
            raise RuntimeError("orc_create failed")
        if eq == "mhd_glm":
            # interior-point weights [nsub][nq^dim][C] and the [nq][nq+2] differentiation matrix
            if cell_w is None or diff_m is None:
                raise ValueError("mhd_glm needs cell_w and diff_m")
            self._cw = np.ascontiguousarray(cell_w, dtype=np.float64)
            self._dm = np.ascontiguousarray(diff_m, dtype=np.float64)
            nsub, C = (8, 33) if dim == 3 else (6, 13)
            assert self._cw.size == nsub * nq ** dim * C and self._dm.size == nq * (nq + 2)
            self.L.orc_set_mhd_tables(self.h, _dp(self._cw), _dp(self._dm))

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def set_ic(self, problem, user=(0.0,) * 8, seed=0):
        u = np.zeros(8)
        u[: len(user)] = user
        if self.L.orc_set_ic(self.h, problem.encode(), _dp(u), int(seed)) != 0:
            raise ValueError(f"unknown problem {problem}")

    @property
    def state(self):
        a = np.empty(self.shape)
        self.L.orc_get_state(self.h, _dp(a))
        return a

    @state.setter
    def state(self, U):
        a = np.ascontiguousarray(U, dtype=np.float64).reshape(self.shape)
        self.L.orc_set_state(self.h, _dp(a))

    @property
    def rhs(self):
        a = np.empty(self.shape)
        self.L.orc_get_rhs(self.h, _dp(a))
        return a

    @property
    def flags(self):
        a = np.empty((2,) + self.shape[1:])
        self.L.orc_get_flags(self.h, _dp(a))
        return a

    @flags.setter
    def flags(self, F):
        a = np.ascontiguousarray(F, dtype=np.float64).reshape((2,) + self.shape[1:])
        self.L.orc_set_flags(self.h, _dp(a))

    def face_states(self, weno=False):
        """Own face states [F*Q, NV, nz, ny, nx]: linear, or full WENO + limiter."""
        nc = int(np.prod(self.shape[1:]))
        F = 2 * self.cfg.dim
        Q = self.cfg.nq ** (self.cfg.dim - 1)
        a = np.empty((F * Q, self.nv) + self.shape[1:])
        assert a.size == F * Q * self.nv * nc
        self.L.orc_face_states(self.h, 1 if weno else 0, _dp(a))
        return a

    def face_states_limited(self):
        """Own face states after the linear reconstruction and the positivity
        limiter, [F*Q, NV, nz, ny, nx] (the limiter's theta is not kept)."""
        nc = int(np.prod(self.shape[1:]))
        F = 2 * self.cfg.dim
        Q = self.cfg.nq ** (self.cfg.dim - 1)
        a = np.empty((F * Q, self.nv) + self.shape[1:])
        assert a.size == F * Q * self.nv * nc
        self.L.orc_face_states(self.h, 2, _dp(a))
        return a

    def weno_omegas(self, i, j, k=0):
        """Nonlinear WENO-AO weights [nSub, NV] of cell (i, j, k)."""
        nsub = 8 if self.cfg.dim == 3 else 6
        assert 0 <= i < self.shape[3] and 0 <= j < self.shape[2] and 0 <= k < self.shape[1]
        a = np.empty((nsub, self.nv))
        self.L.orc_weno_omegas(self.h, i, j, k, _dp(a))
        return a

    def eval_rhs(self, which=0):
        return self.L.orc_eval_rhs(self.h, which)

    def prime(self):
        self.L.orc_prime(self.h)

    def take_step(self):
        return self.L.orc_take_step(self.h)

    def solve(self):
        n = 0
        while self.time[2] < self.cfg.maxSteps and not self.L.orc_last_step(self.h):
            self.take_step()
            n += 1
        return n

    @property
    def time(self):
        t, dt, e = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_int64()
        self.L.orc_get_time(self.h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n), ctypes.byref(e))
        return t.value, dt.value, n.value

    def set_time(self, t, dt):
        self.L.orc_set_time(self.h, t, dt)

    def force_full_weno(self):
        self.L.orc_force_full_weno(self.h)

    @property
    def nweno(self):
        return self.L.orc_num_weno(self.h)

    @property
    def rhs_evals(self):
        return self.L.orc_rhs_evals(self.h)
