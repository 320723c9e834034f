"""weno_solver_noalloc, the allocation-free variant of weno_solver: through
the C ABI, on device buffers, against weno_solver on the same input (bitwise),
twice in a row with the second grid smaller (the shared device scratch keeps
the first call's contents), plus the size limit."""
import ctypes as C

import numpy as np
import pytest

from pypde_amd import _capi, systems
from pypde_amd.solvers import Solver, weno_solver


def _lib():
    L = _capi.lib()
    L.pypde_dev_alloc.restype = C.c_void_p
    L.pypde_dev_alloc.argtypes = [C.c_size_t]
    L.pypde_dev_free.argtypes = [C.c_void_p]
    L.pypde_memcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return L


def _noalloc(L, ctx, u, N):
    u = np.ascontiguousarray(u, dtype=np.float64)
    nX = np.array(u.shape[:-1], dtype=np.int32)
    ndim, V = len(nX), u.shape[-1]
    shp = tuple(int(n) - 2 * (N - 1) for n in nX) + (N,) * ndim + (V,)
    out = np.full(shp, np.nan)
    du = L.pypde_dev_alloc(u.nbytes)
    dr = L.pypde_dev_alloc(out.nbytes)
    assert du and dr
    try:
        _capi.check(L.pypde_memcpy(du, u.ctypes.data, u.nbytes, 1), "memcpy")
        _capi.check(L.weno_solver_noalloc(ctx, dr, du, _capi.ip(nX), None), "weno_solver_noalloc")
        _capi.check(L.pypde_memcpy(out.ctypes.data, dr, out.nbytes, 2), "memcpy")
    finally:
        L.pypde_dev_free(du)
        L.pypde_dev_free(dr)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("name,ndim,N,V,shapes", [
    ("reactive2d", 2, 3, 5, [(41, 37), (23, 19)]),
    ("euler2d", 2, 4, 4, [(50, 44), (12, 30)]),
    ("euler1d", 1, 3, 3, [(300,), (17,)]),
])
def test_noalloc_matches_weno_solver_twice(name, ndim, N, V, shapes):
    L = _lib()
    rng = np.random.default_rng(11 + N + ndim)
    small = np.ones(tuple([2 * N] * ndim) + (V,))
    sol = Solver(small, [1.0] * ndim, systems.source(name), order=N,
                 stiff=(name == "reactive2d"))
    try:
        for shp in shapes:                      # second call: a smaller grid
            u = rng.uniform(0.5, 2.0, shp + (V,))
            got = _noalloc(L, sol.h, u, N)
            want = weno_solver(u, order=N)
            assert got.shape == want.shape
            assert np.array_equal(got, want), shp
    finally:
        sol.close()


@pytest.mark.gpu
def test_noalloc_rejects_grids_beyond_its_scratch():
    L = _lib()
    sol = Solver(np.ones((6, 6, 5)), [1.0, 1.0], systems.source("reactive2d"), order=3, stiff=True)
    try:
        nX = np.array([2048, 2048], dtype=np.int32)   # 4.2M cells x 50 doubles > 64 MiB
        du = L.pypde_dev_alloc(8)
        try:
            rc = L.weno_solver_noalloc(sol.h, du, du, _capi.ip(nX), None)
        finally:
            L.pypde_dev_free(du)
        assert rc == 1 and "limit" in _capi.last_error()
    finally:
        sol.close()
