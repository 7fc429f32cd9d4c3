"""kfvm_eval_rhs_noalloc, the RHS evaluation that allocates no device memory
(every scratch buffer in the library's fixed device array), against
kfvm_eval_rhs on the same state: bit for bit, on two calls in a row where
the second has a smaller grid, for the Euler (KXRCF-sparse, full WENO),
GLM-MHD and linear-advection pipelines.  Device buffers for the state and the
result come from hipMalloc of the HIP runtime the library itself runs on.
"""
import ctypes

import numpy as np
import pytest

from helpers import config, init_text

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the library's own HIP runtime."""

    def __init__(self):
        import kfvm_amd
        kfvm_amd.load_library()
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p.value

    def upload(self, ptr, a):
        assert self.h.hipMemcpy(ptr, a.ctypes.data, a.nbytes, 1) == 0  # host to device

    def download(self, a, ptr):
        assert self.h.hipMemcpy(a.ctypes.data, ptr, a.nbytes, 2) == 0  # device to host

    def free(self, ptr):
        assert self.h.hipFree(ptr) == 0


CASES = {
    "euler3d_sparse": (config(n=(24, 24, 24), wenothresh=1.0, fixeddt=0.004),
                       config(n=(16, 16, 16), wenothresh=1.0, fixeddt=0.004)),
    "euler3d_full": (config(n=(20, 20, 20), wenothresh=-1.0, fixeddt=0.004),
                     config(n=(12, 12, 12), wenothresh=-1.0, fixeddt=0.004)),
    "mhd3d_sparse": (config(n=(16, 16, 16), equation="mhd_glm", riemann="llf", problem="mhd_vortex3d",
                            gamma=5.0 / 3.0, bc="periodic", user=(), wenothresh=1.0, fixeddt=0.002),
                     config(n=(12, 12, 12), equation="mhd_glm", riemann="llf", problem="mhd_vortex3d",
                            gamma=5.0 / 3.0, bc="periodic", user=(), wenothresh=1.0, fixeddt=0.002)),
    "linadv2d_full": (config(dim=2, n=(48, 48, 1), nq=3, equation="linadv", riemann="exact",
                             problem="linadv_sine", user=(1.0, 1.0, 0.0), bc="periodic", wenothresh=-1.0,
                             advvel=(1.0, 1.0, 0.0), fixeddt=0.005),
                      config(dim=2, n=(32, 32, 1), nq=3, equation="linadv", riemann="exact",
                             problem="linadv_sine", user=(1.0, 1.0, 0.0), bc="periodic", wenothresh=-1.0,
                             advvel=(1.0, 1.0, 0.0), fixeddt=0.005)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_noalloc_rhs_equals_eval_rhs(name):
    import kfvm_amd
    hip = _Hip()
    for c in CASES[name]:  # the second call has the smaller grid
        s = kfvm_amd.Solver(init_text=init_text(c))
        assert s.take_step() == 0  # a developed state with limiter thetas from a previous RHS
        U = s.state
        du, dr = hip.alloc(U.nbytes), hip.alloc(U.nbytes)
        try:
            hip.upload(du, U)
            hip.upload(dr, np.full_like(U, np.nan))
            v1 = s.eval_rhs_noalloc(du, dr)
            got = np.empty_like(U)
            hip.download(got, dr)
            # the solver is unchanged: its own evaluation of the same state agrees
            assert np.array_equal(s.state, U)
            v0 = s.eval_rhs(0)
            want = s.rhs
            d = float(np.abs(got - want).max())
            print(f"[noalloc] {name} {U.shape}: max |noalloc - eval_rhs| {d:.3e}, speeds {v1!r} / {v0!r}",
                  flush=True)
            assert np.array_equal(got, want)
            assert v1 == v0
        finally:
            hip.free(du)
            hip.free(dr)
            s.close()


def test_noalloc_rejects_grids_beyond_its_scratch():
    """A 3D Euler grid whose scratch exceeds the fixed array is refused with a
    clear error before any kernel runs, and host pointers are refused."""
    import kfvm_amd
    hip = _Hip()
    s = kfvm_amd.Solver(init_text=init_text(config(n=(64, 64, 64), wenothresh=1.0)))
    U = s.state
    du, dr = hip.alloc(U.nbytes), hip.alloc(U.nbytes)
    try:
        with pytest.raises(RuntimeError, match="bytes of scratch"):
            s.eval_rhs_noalloc(du, dr)
        host = np.zeros_like(U)
        with pytest.raises(RuntimeError, match="device memory"):
            s.eval_rhs_noalloc(host.ctypes.data, dr)
    finally:
        hip.free(du)
        hip.free(dr)
        s.close()
