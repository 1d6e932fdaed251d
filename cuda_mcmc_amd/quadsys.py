"""Batched systems of quadratic equations: MH-RE sampling + Gauss-Newton polish.

The reference's solvsys example (``examples/solvsys/solvsys.py``,
``examples/solvsys/model.cuh:23-57``) samples one 2-equation system through
the generic ``model()`` path.  Here every system is one fit of a session on
the native quadratic-system kernel (MHRE_MODEL_QUADSYS, one lane per chain):
chi^2 = sum_e (dat_e - x^T Q_e x - b_e . x)^2 std2r_e with dat_e = -c_e.
After sampling, the best chain of each system is polished by Gauss-Newton
(Levenberg-Marquardt) steps whose residuals and Jacobians come from the
device kernel behind ``mhre_quad_jacobian``; the n x n solves run on the host
in float64.
"""
import ctypes as C

import numpy as np

from . import _lib
from .engine import Fit, Session


def quad_fit(Q, b, c, pmin=-5.0, pmax=5.0, nbeta=16, nseq=32, nburn=3000, niter=0, seed=0,
             sigma=0.01, record=None, **kw):
    """Fit of one system Q [neq, n, n], b [neq, n], c [neq]."""
    Q = np.asarray(Q, np.float32)
    b = np.asarray(b, np.float32)
    neq, n = b.shape
    coor = np.concatenate([Q.ravel(), b.ravel()]).astype(np.float32)
    dat = (-np.asarray(c, np.float64)).astype(np.float32)
    std2r = np.full(neq, 1.0 / sigma ** 2, np.float32)
    pdescr = np.ones(n, np.int32)
    lo = np.full(n, pmin, np.float32)
    hi = np.full(n, pmax, np.float32)
    if record is None:
        record = niter > 0
    return Fit(pdescr, lo, hi, coor, dat, std2r, nbeta=nbeta, nseq=nseq, nburn=nburn, niter=niter,
               model="quadsys", seed=seed, record=record, **kw)


def best_rows(fits):
    """Lowest-chi^2 chain state of each fit: X [nfits, n], chi^2 [nfits]."""
    X, c2 = [], []
    for f in fits:
        i = np.unravel_index(np.argmin(f.chi2c), f.chi2c.shape)
        X.append(f.pcur[i])
        c2.append(f.chi2c[i])
    return np.array(X, np.float32), np.array(c2, np.float64)


def best_sequence_rows(fits, k=4):
    """Starting points from independent sequences: the lowest-chi^2 chain of
    each of the k sequences whose best chain is lowest, per fit.  Returns
    X [nfits * k, n], chi^2 [nfits * k] and the fit of every row."""
    X, c2, sys = [], [], []
    for s, f in enumerate(fits):
        ib = np.argmin(f.chi2c, axis=0)                   # best temperature slot of every sequence
        best = f.chi2c[ib, np.arange(f.nseq)]
        for q in np.argsort(best, kind="stable")[:k]:
            X.append(f.pcur[ib[q], q])
            c2.append(best[q])
            sys.append(s)
    return np.array(X, np.float32), np.array(c2, np.float64), np.array(sys, np.int32)


def polish_best(Q, b, dat, X, sys, nsys, iters=20, lib=None):
    """Gauss-Newton from every row of X (row i belongs to system sys[i]);
    returns the lowest-residual result of each system: X [nsys, n], |r| [nsys]."""
    Xp, nr = gauss_newton(Q, b, dat, X, iters=iters, sys=sys, lib=lib)
    best = np.full(nsys, np.inf)
    out = np.zeros((nsys, X.shape[1]), np.float32)
    for i in range(len(sys)):
        if nr[i] < best[sys[i]]:
            best[sys[i]] = nr[i]
            out[sys[i]] = Xp[i]
    return out, best


def jacobian(Q, b, dat, X, sys=None, lib=None):
    """Residuals r [rows, neq] = x^T Q_e x + b_e . x - dat_e and J [rows, neq, n] on the GPU."""
    L = _lib.load(lib)
    Q = np.ascontiguousarray(Q, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    dat = np.ascontiguousarray(dat, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    nsys, neq, n = b.shape
    rows = X.shape[0]
    if X.shape[1] != n or Q.shape != (nsys, neq, n, n) or dat.shape != (nsys, neq):
        raise ValueError("inconsistent quadratic-system shapes")
    sysa = None if sys is None else np.ascontiguousarray(sys, np.int32)
    r = np.empty((rows, neq), np.float32)
    J = np.empty((rows, neq, n), np.float32)
    F, I = C.c_float, C.c_int
    rc = L.mhre_quad_jacobian(_lib.ptr(Q, F), _lib.ptr(b, F), _lib.ptr(dat, F), _lib.ptr(sysa, I), nsys, neq,
                              n, _lib.ptr(X, F), rows, _lib.ptr(r, F), _lib.ptr(J, F))
    _lib.check(L, rc, "mhre_quad_jacobian")
    return r, J


def gauss_newton(Q, b, dat, X0, iters=20, lam=1e-6, box=(-5.0, 5.0), sys=None, lib=None):
    """Levenberg-Marquardt polish of X0 [rows, n] (row i solves system sys[i],
    or system i when sys is None).

    Returns (X, |r|) with |r| the residual norm of every row (float32 device
    evaluation, float64 accumulation).
    """
    X = np.asarray(X0, np.float64).copy()
    lamv = np.full(X.shape[0], float(lam))
    r, J = jacobian(Q, b, dat, X, sys=sys, lib=lib)
    nr = np.linalg.norm(r.astype(np.float64), axis=1)
    for _ in range(iters):
        Jd = J.astype(np.float64)
        JtJ = np.einsum("sek,sel->skl", Jd, Jd)
        Jtr = np.einsum("sek,se->sk", Jd, r.astype(np.float64))
        A = JtJ + lamv[:, None, None] * np.eye(X.shape[1])[None] * np.maximum(
            np.einsum("skk->sk", JtJ)[:, :, None], 1e-12)
        step = -np.linalg.solve(A, Jtr[..., None])[..., 0]
        trial = np.clip(X + step, box[0], box[1])
        r2, J2 = jacobian(Q, b, dat, trial, sys=sys, lib=lib)
        n2 = np.linalg.norm(r2.astype(np.float64), axis=1)
        ok = n2 < nr
        X[ok], r[ok], J[ok], nr[ok] = trial[ok], r2[ok], J2[ok], n2[ok]
        lamv[ok] = np.maximum(lamv[ok] / 10.0, 1e-12)
        lamv[~ok] = lamv[~ok] * 10.0
    return X.astype(np.float32), nr


def solve_systems(ds, nbeta=16, nseq=32, nburn=3000, niter=0, seed=0, gn_iters=20, sigma=0.01, lib=None,
                  **session_options):
    """Sample every system of ds (quadsys_dataset layout) in one session, then polish.

    Returns dict(X, resid, chi2_best, fits).
    """
    Q, b, c = ds["Q"], ds["b"], ds["c"]
    fits = [quad_fit(Q[s], b[s], c[s], nbeta=nbeta, nseq=nseq, nburn=nburn, niter=niter, seed=seed + s,
                     sigma=sigma) for s in range(len(c))]
    with Session(fits, lib=lib, **session_options) as ses:
        ses.run(nburn + niter)
        ses.download()
    X0, c2 = best_rows(fits)
    X, resid = gauss_newton(Q, b, -np.asarray(c), X0, iters=gn_iters, lib=lib)
    return dict(X=X, resid=resid, chi2_best=c2, fits=fits)
