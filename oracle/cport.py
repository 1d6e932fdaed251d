"""ctypes binding of the C / OpenMP oracle port (oracle/c/mhre_oracle.c).

Test infrastructure and the timed CPU baseline of bench.py only.
Parity: pinned against the NumPy oracle by tests/test_oracle_c.py.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libmhre_oracle.so")

_L = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def load():
    global _L
    if _L is None:
        if not os.path.isfile(LIB):
            build()
        L = C.CDLL(LIB)
        f, i, d = C.POINTER(C.c_float), C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.oracle_eht_chi2_batch.restype = None
        L.oracle_eht_chi2_batch.argtypes = [C.c_int, f, C.c_int, C.c_int, C.c_int, f, f, f, i, d, C.c_int]
        L.oracle_eht_iteration.restype = C.c_longlong
        L.oracle_eht_iteration.argtypes = ([C.c_int] * 8 + [C.c_uint64, C.c_float, i, i, i, f, f, f,
                                            C.c_int, C.c_int, f, f, f, i,
                                            f, f, f, i, i, i, i, i, i, f, i, f, i, f, C.c_int, C.c_int])
        _L = L
    return _L


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def eht_chi2_batch(P, d, nthreads=0):
    L = load()
    P = np.ascontiguousarray(P, np.float32).reshape(-1, P.shape[-1])
    out = np.zeros(P.shape[0])
    coor = np.ascontiguousarray(d["coor"], np.float32)
    dat = np.ascontiguousarray(d["dat"], np.float32)
    w = np.ascontiguousarray(d["std2r"], np.float32)
    ic = np.ascontiguousarray(d["icoor"], np.int32)
    L.oracle_eht_chi2_batch(P.shape[0], _p(P, C.c_float), P.shape[1], int(d["nvis"]), int(d["ncph"]),
                            _p(coor, C.c_float), _p(dat, C.c_float), _p(w, C.c_float), _p(ic, C.c_int),
                            _p(out, C.c_double), int(nthreads))
    return out


def eht_iteration(fit, d, it, nthreads=0):
    """One MH-RE iteration on a cuda_mcmc_amd.engine.Fit-like object's arrays (in place)."""
    L = load()
    burn = it < fit.nburn
    itr = it - fit.nburn
    coor = np.ascontiguousarray(d["coor"], np.float32)
    dat = np.ascontiguousarray(d["dat"], np.float32)
    w = np.ascontiguousarray(d["std2r"], np.float32)
    ic = np.ascontiguousarray(d["icoor"], np.int32)
    F, I = C.c_float, C.c_int
    return L.oracle_eht_iteration(
        fit.nbeta, fit.nseq, fit.nptot, fit.nprm, fit.nadj, int(it), int(burn), int(itr),
        fit.seed, fit.adapt_target,
        _p(fit.pdescr, I), _p(fit.ivar, I), _p(fit.invar, I), _p(fit.pmint, F), _p(fit.pmaxt, F), _p(fit.beta, F),
        int(d["nvis"]), int(d["ncph"]), _p(coor, F), _p(dat, F), _p(w, F), _p(ic, I),
        _p(fit.pcur, F), _p(fit.chi2c, F), _p(fit.pstp, F), _p(fit.n_cnt, I), _p(fit.n_acpt, I),
        _p(fit.n_hist, I), _p(fit.n_exch, I), _p(fit.tcur, I), _p(fit.flag, I), _p(fit.ptent, F),
        _p(fit.ptentn, I), _p(fit.pout if fit.niter else None, F), _p(fit.tout if fit.niter else None, I),
        _p(fit.chi2 if fit.niter else None, F), max(fit.niter, 1), int(nthreads))
