"""cs_solve_bsr_noalloc (the solve that allocates nothing) against cs_solve, through the C-ABI.

Two calls in a row on one GPU, the second on a smaller system, each on the system a context
assembled (hang33, then hang17, four steps in): the result is within what cs_solve's own
tolerance allows (positions x = x0 + dt (v + dv) within 1e-5 m of cs_solve's, so |dv| within
1e-5 / dt), meets the 1e-5 residual in FP64 on the host, and takes the same iteration count
within two (the same algorithm; only the summation order differs).  The device buffers are
made by the test (hipMalloc through the library's own HIP runtime); the entry point itself
allocates nothing.
"""
import ctypes as C

import numpy as np
import pytest

from clothsimulator_amd import Simulator, load_library, scenes
from clothsimulator_amd._cabi import CS_E_ARG, CS_E_CAPACITY, CS_OK

pytestmark = pytest.mark.gpu

MAX_NODES = 65536  # CS_NOALLOC_MAX_NODES


def _hip():
    """The HIP runtime the library itself runs on (already loaded by it)."""
    H = C.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipFree.argtypes = [C.c_void_p]
    return H


class DeviceArray:
    """A device copy of a host array, made and freed by the test (hipMalloc / hipMemcpy)."""

    def __init__(self, host):
        self.H = _hip()
        self.host = np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert self.H.hipMalloc(C.byref(self.ptr), max(self.host.nbytes, 4)) == 0
        assert self.H.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1) == 0  # host to device

    def get(self):
        out = np.empty_like(self.host)
        assert self.H.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # device to host
        return out

    def __del__(self):
        self.H.hipFree(self.ptr)


def _noalloc(rp, ci, vals, b, tol=1e-5, maxit=4000):
    L = load_library()
    drp, dci = DeviceArray(rp.astype(np.int32)), DeviceArray(ci.astype(np.int32))
    dvals, db = DeviceArray(vals.reshape(-1).astype(np.float32)), DeviceArray(b.astype(np.float32))
    dx = DeviceArray(np.full(len(b), np.nan, np.float32))
    it, rel = C.c_int(-1), C.c_float(-1.0)
    st = L.cs_solve_bsr_noalloc(None, len(rp) - 1, drp.ptr, dci.ptr, dvals.ptr, db.ptr, dx.ptr, tol, maxit,
                                C.byref(it), C.byref(rel))
    return st, dx.get(), it.value, rel.value


def test_noalloc_solve_matches_cs_solve_on_two_calls():
    import scipy.sparse as sp
    for n in (33, 17):  # the second call on a smaller system
        spec = scenes.hang(n)
        sim = Simulator(spec)
        s, _ = sim.run(4)
        assert s == CS_OK
        rp, ci, vals, b = sim.assemble(0)
        dv, st = sim.solve(0)
        status, x, iters, rel = _noalloc(rp, ci, vals, b)
        A = sp.bsr_matrix((vals.astype(np.float64), ci, rp), shape=(len(b), len(b)))
        res = np.linalg.norm(b.astype(np.float64) - A @ x.astype(np.float64)) / np.linalg.norm(b)
        dt = spec.params.dt
        dx = dt * float(np.abs(x - dv).max())
        print("hang%d: cs_solve %d iterations, no-alloc %d (status %d, recursive %.3g, FP64 %.3g); "
              "positions differ by %.3g m" % (n, st.pcg_iters, iters, status, rel, res, dx))
        assert status == CS_OK
        assert np.isfinite(x).all()
        assert res <= 2e-5 and rel <= 1e-5
        # cs_solve confirms its stop on the true residual and, when that misses, continues a little
        # below the tolerance; the no-alloc solve stops on its recursive residual
        assert iters <= st.pcg_iters + 2 and st.pcg_iters <= max(iters + 2, 1.1 * iters)
        assert dx <= 1e-5
        sim.close()


def test_noalloc_rejects_oversize_and_malformed_systems():
    L = load_library()
    it = C.c_int()
    # beyond the fixed scratch: rejected before any pointer is read
    assert L.cs_solve_bsr_noalloc(None, MAX_NODES + 1, None, None, None, None, None, 1e-5, 10, C.byref(it),
                                  None) == CS_E_CAPACITY
    assert L.cs_solve_bsr_noalloc(None, 0, None, None, None, None, None, 1e-5, 10, None, None) == CS_E_ARG
    # a row without its diagonal block, and a column out of range
    rp = np.array([0, 1, 2], np.int32)
    b = np.ones(6, np.float32)
    vals = np.tile(np.eye(3, dtype=np.float32).reshape(1, 9), (2, 1))
    for ci in (np.array([1, 0], np.int32), np.array([0, 5], np.int32)):
        status, _, _, _ = _noalloc(rp, ci, vals, b)
        assert status == CS_E_ARG
    # the same arrays with valid columns solve (identity system: one iteration)
    status, x, iters, _ = _noalloc(rp, np.array([0, 1], np.int32), vals, b)
    assert status == CS_OK and iters == 1 and np.allclose(x, 1.0)
