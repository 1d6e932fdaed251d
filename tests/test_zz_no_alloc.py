"""mhre_quad_jacobian_noalloc (stages through the library's static device
area) against mhre_quad_jacobian (a per-call allocation): identical residuals
and Jacobians on two calls in a row, the second smaller (the area keeps the
first call's contents), and a clear error above the area's size."""
import ctypes as C

import numpy as np
import pytest

from cuda_mcmc_amd import _lib

pytestmark = pytest.mark.gpu

fp = C.POINTER(C.c_float)
ip = C.POINTER(C.c_int)


def _call(L, fn, Q, b, dat, sys, X):
    nsys, neq, n = b.shape
    rows = X.shape[0]
    r = np.full((rows, neq), np.nan, np.float32)
    J = np.full((rows, neq, n), np.nan, np.float32)
    rc = getattr(L, fn)(Q.ctypes.data_as(fp), b.ctypes.data_as(fp), dat.ctypes.data_as(fp),
                        None if sys is None else sys.ctypes.data_as(ip), nsys, neq, n,
                        X.ctypes.data_as(fp), rows, r.ctypes.data_as(fp), J.ctypes.data_as(fp))
    return rc, r, J


def _systems(rng, nsys, neq, n, rows, with_sys):
    Q = rng.normal(size=(nsys, neq, n, n)).astype(np.float32)
    b = rng.normal(size=(nsys, neq, n)).astype(np.float32)
    dat = rng.normal(size=(nsys, neq)).astype(np.float32)
    X = rng.uniform(-2, 2, size=(rows, n)).astype(np.float32)
    sys = rng.integers(0, nsys, size=rows).astype(np.int32) if with_sys else None
    return Q, b, dat, sys, X


def test_noalloc_matches_allocating_entry_point():
    L = _lib.load()
    rng = np.random.default_rng(0)
    for nsys, neq, n, rows, with_sys in [(600, 8, 8, 1500, True), (40, 3, 4, 40, False)]:
        Q, b, dat, sys, X = _systems(rng, nsys, neq, n, rows, with_sys)
        rc_a, r_a, J_a = _call(L, "mhre_quad_jacobian", Q, b, dat, sys, X)
        rc_b, r_b, J_b = _call(L, "mhre_quad_jacobian_noalloc", Q, b, dat, sys, X)
        assert rc_a == 0 and rc_b == 0, L.mhre_last_error()
        np.testing.assert_array_equal(r_b, r_a)
        np.testing.assert_array_equal(J_b, J_a)
        # the residual against a float64 evaluation on the host
        s = sys if with_sys else np.arange(rows)
        x = X.astype(np.float64)
        lhs = np.einsum("ri,reij,rj->re", x, Q[s].astype(np.float64), x) + np.einsum("rei,ri->re", b[s], x)
        np.testing.assert_allclose(r_b, lhs - dat[s], rtol=1e-4, atol=1e-4)


def test_noalloc_rejects_calls_beyond_the_area():
    L = _lib.load()
    nsys, neq, n = 20000, 8, 8            # Q alone is 40 MiB
    Q = np.zeros((nsys, neq, n, n), np.float32)
    b = np.zeros((nsys, neq, n), np.float32)
    dat = np.zeros((nsys, neq), np.float32)
    X = np.zeros((4, n), np.float32)
    rc, _, _ = _call(L, "mhre_quad_jacobian_noalloc", Q, b, dat, None, X)
    assert rc == -1 and b"static scratch" in L.mhre_last_error()
