"""Python mirror of the reference's Solver surface over the C ABI of
libkfvm_amd.so (include/kfvm.h).

    from kfvm_amd import Solver
    s = Solver(init_text=open("riem.init").read())
    s.solve()
    U = s.state            # numpy [NV, nz, ny, nx]

Method names follow KFVM::Solver (src/Solver.H:54-160): solve, take_step,
eval_rhs.  The compute path is the HIP library only; importing this module on
a machine without the built library or without a GPU raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "build", "libkfvm_amd.so"))

ACCEPTED, TOLERANCE, UNPHYSICAL, REJECT_LIMIT, STAGNATED = range(5)
STATUS_NAMES = {ACCEPTED: "accepted", TOLERANCE: "tolerance", UNPHYSICAL: "unphysical",
                REJECT_LIMIT: "reject_limit", STAGNATED: "stagnated"}

# Every extern "C" symbol of include/kfvm.h.
EXPORTS = [
    "kfvm_create", "kfvm_create_from_text", "kfvm_destroy", "kfvm_solve", "kfvm_take_step",
    "kfvm_eval_rhs", "kfvm_advance_stages", "kfvm_get_state", "kfvm_set_state", "kfvm_get_rhs", "kfvm_get_flags",
    "kfvm_set_flags", "kfvm_get_time", "kfvm_set_time", "kfvm_get_info", "kfvm_get_counters",
    "kfvm_force_full_weno", "kfvm_get_weights", "kfvm_set_weights", "kfvm_profile_enable",
    "kfvm_profile_get", "kfvm_profile_reset", "kfvm_sync", "kfvm_weight_sizes",
    "kfvm_build_weights", "kfvm_build_cell_weights", "kfvm_diff_matrix", "kfvm_parse_init",
    "kfvm_last_error", "kfvm_set_indicator", "kfvm_get_indicator_counters", "kfvm_indicator_create",
    "kfvm_indicator_destroy", "kfvm_indicator_collect", "kfvm_indicator_samples", "kfvm_indicator_init",
    "kfvm_indicator_train", "kfvm_indicator_get_params", "kfvm_indicator_set_params",
    "kfvm_indicator_get_weights", "kfvm_indicator_gradient", "kfvm_indicator_logits",
    "kfvm_indicator_sizes", "kfvm_set_nad", "kfvm_write_plot", "kfvm_write_checkpoint", "kfvm_flow_stats",
    "kfvm_write_output_file", "kfvm_read_checkpoint_file", "kfvm_create_batch", "kfvm_get_slice_times",
    "kfvm_initial_condition", "kfvm_eval_rhs_noalloc",
]


class KfvmInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "dim", "nq", "nv", "nx", "ny", "nz", "ncell", "nsub", "nfaces", "nqface",
        "ncells_stencil", "nindic", "rk_stages")] + [
        (n, ctypes.c_double) for n in ("dx", "dy", "dz", "gamma")]


_lib = None


def load_library(path=None):
    """Load libkfvm_amd.so and declare the C signatures (raises if absent)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libkfvm_amd.so not built ({p}); run kfvm-kokkos_amd/build.py")
    lib = ctypes.CDLL(p)
    P = ctypes.c_void_p
    D = ctypes.POINTER(ctypes.c_double)
    I64 = ctypes.c_int64
    F32 = ctypes.POINTER(ctypes.c_float)
    U16 = ctypes.POINTER(ctypes.c_uint16)
    sig = {
        "kfvm_create": [ctypes.c_char_p, ctypes.POINTER(P)],
        "kfvm_create_from_text": [ctypes.c_char_p, ctypes.POINTER(P)],
        "kfvm_destroy": [P],
        "kfvm_solve": [P, ctypes.POINTER(I64)],
        "kfvm_take_step": [P, ctypes.POINTER(ctypes.c_int)],
        "kfvm_eval_rhs": [P, ctypes.c_int, ctypes.c_double, D],
        "kfvm_eval_rhs_noalloc": [P, P, P, ctypes.c_double, D],
        "kfvm_advance_stages": [P, I64, ctypes.c_double],
        "kfvm_get_state": [P, D, I64],
        "kfvm_set_state": [P, D, I64],
        "kfvm_get_rhs": [P, D, I64],
        "kfvm_get_flags": [P, D, I64],
        "kfvm_set_flags": [P, D, I64],
        "kfvm_get_time": [P, D, D, ctypes.POINTER(I64)],
        "kfvm_set_time": [P, ctypes.c_double, ctypes.c_double],
        "kfvm_get_info": [P, ctypes.POINTER(KfvmInfo)],
        "kfvm_get_counters": [P] + [ctypes.POINTER(I64)] * 4,
        "kfvm_force_full_weno": [P],
        "kfvm_get_weights": [P, D, I64, D, I64],
        "kfvm_set_weights": [P, D, I64, D, I64],
        "kfvm_profile_enable": [P, ctypes.c_int],
        "kfvm_profile_get": [P, ctypes.c_char_p, D, ctypes.POINTER(I64)],
        "kfvm_profile_reset": [P],
        "kfvm_sync": [P],
        "kfvm_weight_sizes": [ctypes.c_int, ctypes.c_int, ctypes.POINTER(I64), ctypes.POINTER(I64)],
        "kfvm_build_weights": [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4 + [D, I64, D, I64],
        "kfvm_build_cell_weights": [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4 + [D, I64],
        "kfvm_diff_matrix": [ctypes.c_int, D, I64],
        "kfvm_parse_init": [ctypes.c_char_p, ctypes.c_char_p, I64],
        "kfvm_initial_condition": [ctypes.c_char_p, D, I64],
        "kfvm_set_indicator": [P, U16, I64, ctypes.c_double],
        "kfvm_get_indicator_counters": [P, ctypes.POINTER(I64), ctypes.POINTER(I64), ctypes.POINTER(ctypes.c_int)],
        "kfvm_set_nad": [P, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double],
        "kfvm_create_batch": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64), I64, ctypes.POINTER(P)],
        "kfvm_get_slice_times": [P, D, D, I64],
        "kfvm_write_plot": [P, ctypes.c_char_p, I64],
        "kfvm_write_checkpoint": [P, ctypes.c_char_p, I64],
        "kfvm_flow_stats": [P, D, ctypes.POINTER(ctypes.c_int)],
        "kfvm_write_output_file": [ctypes.c_char_p, ctypes.c_int, D, D, I64, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_char_p, I64],
        "kfvm_read_checkpoint_file": [ctypes.c_char_p, ctypes.c_char_p, D, I64, ctypes.POINTER(ctypes.c_int),
                                      D, D, D],
        "kfvm_indicator_create": [ctypes.c_int, ctypes.POINTER(P)],
        "kfvm_indicator_destroy": [P],
        "kfvm_indicator_collect": [P, P, ctypes.c_double, ctypes.POINTER(I64), ctypes.POINTER(I64)],
        "kfvm_indicator_samples": [P, F32, F32, I64, I64],
        "kfvm_indicator_init": [P, ctypes.c_uint64],
        "kfvm_indicator_train": [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)],
        "kfvm_indicator_get_params": [P, F32, I64],
        "kfvm_indicator_set_params": [P, F32, I64],
        "kfvm_indicator_get_weights": [P, U16, I64],
        "kfvm_indicator_gradient": [P, F32, F32, F32, I64, ctypes.c_double, F32,
                                    ctypes.POINTER(ctypes.c_double)],
        "kfvm_indicator_logits": [P, F32, F32, I64, F32],
        "kfvm_indicator_sizes": [ctypes.c_int, ctypes.POINTER(I64), ctypes.POINTER(I64)],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.kfvm_last_error.argtypes = []
    lib.kfvm_last_error.restype = ctypes.c_char_p
    if path is None:
        _lib = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class KfvmError(RuntimeError):
    pass


def _check(lib, rc, name):
    if rc != 0:
        raise KfvmError(f"{name}: {lib.kfvm_last_error().decode()}")


def build_weights(dim, nq, lfac, dx, dy, dz=1.0):
    """Host-side GP weight tables (face [nsub,2*dim,Q,C], deriv [nsub,I,C])."""
    lib = load_library()
    nf, nd = ctypes.c_int64(), ctypes.c_int64()
    _check(lib, lib.kfvm_weight_sizes(dim, nq, ctypes.byref(nf), ctypes.byref(nd)), "kfvm_weight_sizes")
    nsub, C, I = (8, 33, 19) if dim == 3 else (6, 13, 5)
    Q = nq * nq if dim == 3 else nq
    f = np.empty((nsub, 2 * dim, Q, C))
    d = np.empty((nsub, I, C))
    assert f.size == nf.value and d.size == nd.value
    _check(lib, lib.kfvm_build_weights(dim, nq, lfac, dx, dy, dz, _dptr(f), f.size, _dptr(d), d.size),
           "kfvm_build_weights")
    return f, d


def build_cell_weights(dim, nq, lfac, dx, dy, dz=1.0):
    """Interior-point GP weights [nsub, nq^dim, C] (point (a*nq+b)*nq+c)."""
    lib = load_library()
    nsub, C = (8, 33) if dim == 3 else (6, 13)
    w = np.empty((nsub, nq ** dim, C))
    _check(lib, lib.kfvm_build_cell_weights(dim, nq, lfac, dx, dy, dz, _dptr(w), w.size),
           "kfvm_build_cell_weights")
    return w


def diff_matrix(nq):
    """Lagrange differentiation matrix [nq, nq+2] on {-1/2, Gauss points, +1/2}."""
    lib = load_library()
    m = np.empty((nq, nq + 2))
    _check(lib, lib.kfvm_diff_matrix(nq, _dptr(m), m.size), "kfvm_diff_matrix")
    return m


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _u16ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))


def indicator_sizes(dim):
    """(features per cell, network parameters) of the learned indicator."""
    lib = load_library()
    ni, npar = ctypes.c_int64(), ctypes.c_int64()
    _check(lib, lib.kfvm_indicator_sizes(dim, ctypes.byref(ni), ctypes.byref(npar)), "kfvm_indicator_sizes")
    return ni.value, npar.value


def bf16_to_float(w):
    """bfloat16 bit patterns (uint16) -> float32 values."""
    return (np.asarray(w, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def float_to_bf16(p):
    """float32 -> bfloat16 bit patterns, round to nearest even (finite inputs)."""
    u = np.ascontiguousarray(p, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


class Indicator:
    """Training set and trainer of the learned troubled-cell indicator (GPU)."""

    def __init__(self, dim):
        self._lib = load_library()
        self._h = ctypes.c_void_p()
        _check(self._lib, self._lib.kfvm_indicator_create(dim, ctypes.byref(self._h)), "kfvm_indicator_create")
        self.dim = dim
        self.nin, self.np = indicator_sizes(dim)
        self.nsamples = self.npositive = 0

    def _call(self, name, *args):
        rc = getattr(self._lib, name)(*args)
        if rc != 0:
            raise KfvmError(f"{name}: {self._lib.kfvm_last_error().decode()}")

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.kfvm_indicator_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def collect(self, solver, tau):
        n, npos = ctypes.c_int64(), ctypes.c_int64()
        self._call("kfvm_indicator_collect", self._h, solver._h, float(tau), ctypes.byref(n), ctypes.byref(npos))
        self.nsamples, self.npositive = n.value, npos.value
        return n.value, npos.value

    def samples(self, offset, n):
        X = np.empty((n, self.nin), dtype=np.float32)
        Y = np.empty(n, dtype=np.float32)
        self._call("kfvm_indicator_samples", self._h, _fptr(X), _fptr(Y), int(offset), int(n))
        return X, Y

    def init(self, seed):
        self._call("kfvm_indicator_init", self._h, int(seed))

    def train(self, iters, batch, lr, pos_weight=1.0, seed=0):
        loss = ctypes.c_double()
        self._call("kfvm_indicator_train", self._h, int(iters), int(batch), float(lr), float(pos_weight),
                   int(seed), ctypes.byref(loss))
        return loss.value

    @property
    def params(self):
        P = np.empty(self.np, dtype=np.float32)
        self._call("kfvm_indicator_get_params", self._h, _fptr(P), P.size)
        return P

    @params.setter
    def params(self, P):
        a = np.ascontiguousarray(P, dtype=np.float32)
        self._call("kfvm_indicator_set_params", self._h, _fptr(a), a.size)

    @property
    def weights(self):
        """bfloat16 parameters (uint16 bit patterns) the forward pass uses."""
        W = np.empty(self.np, dtype=np.uint16)
        self._call("kfvm_indicator_get_weights", self._h, _u16ptr(W), W.size)
        return W

    def gradient(self, P, X, Y, pos_weight=1.0):
        P = np.ascontiguousarray(P, dtype=np.float32)
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        g = np.empty(self.np, dtype=np.float32)
        loss = ctypes.c_double()
        self._call("kfvm_indicator_gradient", self._h, _fptr(P), _fptr(X), _fptr(Y), len(Y), float(pos_weight),
                   _fptr(g), ctypes.byref(loss))
        return g, loss.value

    def logits(self, P, X):
        P = np.ascontiguousarray(P, dtype=np.float32)
        X = np.ascontiguousarray(X, dtype=np.float32)
        z = np.empty(len(X), dtype=np.float32)
        self._call("kfvm_indicator_logits", self._h, _fptr(P), _fptr(X), len(X), _fptr(z))
        return z


def write_output_file(init_text, U, flags, step=0, time=0.0, dt=0.0, err_est=1.0, checkpoint=False):
    """Plot or checkpoint file of host arrays (no solver); returns its path."""
    lib = load_library()
    U = np.ascontiguousarray(U, dtype=np.float64)
    F = np.ascontiguousarray(flags, dtype=np.float64)
    buf = ctypes.create_string_buffer(4096)
    _check(lib, lib.kfvm_write_output_file(init_text.encode(), 1 if checkpoint else 0, _dptr(U), _dptr(F),
                                           U.size, int(step), float(time), float(dt), float(err_est), buf,
                                           len(buf)), "kfvm_write_output_file")
    return buf.value.decode()


def read_checkpoint_file(init_text, path, shape):
    """(U, step, time, dt, errEst) of a checkpoint; shape = (NV, nz, ny, nx)."""
    lib = load_library()
    U = np.empty(shape)
    st = ctypes.c_int()
    t, dt, e = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    _check(lib, lib.kfvm_read_checkpoint_file(init_text.encode(), path.encode(), _dptr(U), U.size,
                                              ctypes.byref(st), ctypes.byref(t), ctypes.byref(dt),
                                              ctypes.byref(e)), "kfvm_read_checkpoint_file")
    return U, st.value, t.value, dt.value, e.value


def initial_condition(text):
    """The cell averages kfvm_create fills for an init text (host only, no GPU):
    an array [NV, nz, ny, nx]."""
    import math
    lib = load_library()
    p = parse_init(text)
    nv = {0: 5, 1: 1, 2: 9}[p["eq"]]
    n = p["n"]
    nz = n[2] if p["dim"] == 3 else 1
    U = np.empty((nv, nz, n[1], n[0]))
    _check(lib, lib.kfvm_initial_condition(text.encode(), _dptr(U), U.size), "kfvm_initial_condition")
    return U


def parse_init(text):
    """The library's reading of an init file, as a dict."""
    import json
    lib = load_library()
    buf = ctypes.create_string_buffer(8192)
    _check(lib, lib.kfvm_parse_init(text.encode(), buf, len(buf)), "kfvm_parse_init")
    return json.loads(buf.value.decode())


class Solver:
    """One problem on one GPU (KFVM::Solver)."""

    def __init__(self, init_path=None, init_text=None, seeds=None):
        """seeds: a batch of len(seeds) independent 2D problems (init_text only)."""
        self._lib = load_library()
        self._h = ctypes.c_void_p()
        if (init_path is None) == (init_text is None):
            raise ValueError("give exactly one of init_path / init_text")
        if seeds is not None:
            if init_text is None:
                raise ValueError("batches take init_text")
            sd = (ctypes.c_uint64 * len(seeds))(*[int(x) for x in seeds])
            self._call("kfvm_create_batch", init_text.encode(), sd, len(seeds), ctypes.byref(self._h))
        elif init_path is not None:
            self._call("kfvm_create", init_path.encode(), ctypes.byref(self._h))
        else:
            self._call("kfvm_create_from_text", init_text.encode(), ctypes.byref(self._h))
        info = KfvmInfo()
        self._call("kfvm_get_info", self._h, ctypes.byref(info))
        self.info = info
        self.shape = (info.nv, info.nz, info.ny, info.nx)

    def _call(self, name, *args):
        rc = getattr(self._lib, name)(*args)
        if rc != 0:
            raise KfvmError(f"{name}: {self._lib.kfvm_last_error().decode()}")

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.kfvm_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- Solver members
    def solve(self):
        n = ctypes.c_int64()
        self._call("kfvm_solve", self._h, ctypes.byref(n))
        return n.value

    def take_step(self):
        st = ctypes.c_int()
        self._call("kfvm_take_step", self._h, ctypes.byref(st))
        return st.value

    def eval_rhs(self, which=0, t=None):
        """Solver::evalRHS of the current (0) or step-start (1) state at time t
        (default: the solver's time)."""
        v = ctypes.c_double()
        if t is None:
            t = self.time[0]
        self._call("kfvm_eval_rhs", self._h, int(which), float(t), ctypes.byref(v))
        return v.value

    def eval_rhs_noalloc(self, u_ptr, rhs_ptr, t=None):
        """The RHS of the device state at address u_ptr ([NV][nz][ny][nx]
        doubles) into the device buffer at rhs_ptr, with no device allocation
        (scratch in the library's fixed device array); the solver is left
        unchanged.  Returns the largest wave speed."""
        v = ctypes.c_double()
        if t is None:
            t = self.time[0]
        self._call("kfvm_eval_rhs_noalloc", self._h, ctypes.c_void_p(u_ptr), ctypes.c_void_p(rhs_ptr), float(t),
                   ctypes.byref(v))
        return v.value

    def advance_stages(self, n, dt):
        """Queue exactly n RK stages of fixed-dt steps (no host sync)."""
        self._call("kfvm_advance_stages", self._h, int(n), float(dt))

    def force_full_weno(self):
        self._call("kfvm_force_full_weno", self._h)

    def set_indicator(self, weights=None, threshold=0.0):
        """Learned WENO selection with bfloat16 weights (uint16); None: KXRCF."""
        if weights is None:
            self._call("kfvm_set_indicator", self._h, None, 0, 0.0)
            return
        w = np.ascontiguousarray(weights, dtype=np.uint16)
        self._call("kfvm_set_indicator", self._h, _u16ptr(w), w.size, float(threshold))

    def write_plot(self):
        buf = ctypes.create_string_buffer(4096)
        self._call("kfvm_write_plot", self._h, buf, len(buf))
        return buf.value.decode()

    def write_checkpoint(self):
        buf = ctypes.create_string_buffer(4096)
        self._call("kfvm_write_checkpoint", self._h, buf, len(buf))
        return buf.value.decode()

    @property
    def flow_stats(self):
        """Flow statistics of the current state (Solver::evalFlowStats): Euler
        (kinetic energy, enstrophy), GLM-MHD (the same and B . curl B), linear
        advection (0,); domain averages."""
        out = np.zeros(3)
        n = ctypes.c_int(0)
        self._call("kfvm_flow_stats", self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                   ctypes.byref(n))
        return out[: n.value].copy()

    def set_nad(self, on=True, range_frac=0.05, value_frac=1e-3, max_frac=0.02):
        """Admissibility check of learned steps: tolerance max(value_frac |bound|,
        range_frac range); a step with more than max_frac failing cells is redone."""
        self._call("kfvm_set_nad", self._h, 1 if on else 0, float(range_frac), float(value_frac),
                   float(max_frac))

    @property
    def indicator_counters(self):
        n, cells, on = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        self._call("kfvm_get_indicator_counters", self._h, ctypes.byref(n), ctypes.byref(cells), ctypes.byref(on))
        return {"nad_fallbacks": n.value, "nad_cells": cells.value, "learned": bool(on.value)}

    # ---- data
    @property
    def state(self):
        a = np.empty(self.shape, dtype=np.float64)
        self._call("kfvm_get_state", self._h, _dptr(a), a.size)
        return a

    @state.setter
    def state(self, U):
        a = np.ascontiguousarray(U, dtype=np.float64).reshape(self.shape)
        self._call("kfvm_set_state", self._h, _dptr(a), a.size)

    @property
    def rhs(self):
        a = np.empty(self.shape, dtype=np.float64)
        self._call("kfvm_get_rhs", self._h, _dptr(a), a.size)
        return a

    @property
    def flags(self):
        a = np.empty((2,) + self.shape[1:], dtype=np.float64)
        self._call("kfvm_get_flags", self._h, _dptr(a), a.size)
        return a

    @flags.setter
    def flags(self, F):
        a = np.ascontiguousarray(F, dtype=np.float64).reshape((2,) + self.shape[1:])
        self._call("kfvm_set_flags", self._h, _dptr(a), a.size)

    @property
    def time(self):
        t, dt, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        self._call("kfvm_get_time", self._h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n))
        return t.value, dt.value, n.value

    @property
    def slice_times(self):
        """(times, step sizes) of the problems of a batch."""
        n = self.shape[1] if self.info.dim == 2 else 1
        t, dt = np.empty(n), np.empty(n)
        self._call("kfvm_get_slice_times", self._h, _dptr(t), _dptr(dt), n)
        return t, dt

    def set_time(self, t, dt):
        self._call("kfvm_set_time", self._h, float(t), float(dt))

    @property
    def counters(self):
        v = [ctypes.c_int64() for _ in range(4)]
        self._call("kfvm_get_counters", self._h, *[ctypes.byref(x) for x in v])
        return dict(zip(("rhs_evals", "rejects_tol", "rejects_unphys", "last_nweno"),
                        [x.value for x in v]))

    def weight_shapes(self):
        i = self.info
        return ((i.nsub, i.nfaces, i.nqface, i.ncells_stencil), (i.nsub, i.nindic, i.ncells_stencil))

    @property
    def weights(self):
        fs, ds = self.weight_shapes()
        f = np.empty(fs)
        d = np.empty(ds)
        self._call("kfvm_get_weights", self._h, _dptr(f), f.size, _dptr(d), d.size)
        return f, d

    def set_weights(self, face, deriv):
        f = np.ascontiguousarray(face, dtype=np.float64)
        d = np.ascontiguousarray(deriv, dtype=np.float64)
        self._call("kfvm_set_weights", self._h, _dptr(f), f.size, _dptr(d), d.size)

    # ---- device timing
    def profile(self, on=True):
        self._call("kfvm_profile_enable", self._h, 1 if on else 0)

    def profile_reset(self):
        self._call("kfvm_profile_reset", self._h)

    def profile_get(self, name):
        ms, n = ctypes.c_double(), ctypes.c_int64()
        self._call("kfvm_profile_get", self._h, name.encode(), ctypes.byref(ms), ctypes.byref(n))
        return ms.value, n.value

    def sync(self):
        self._call("kfvm_sync", self._h)
