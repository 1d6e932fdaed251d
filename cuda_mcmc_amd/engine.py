"""Batched, device-resident MH-RE sessions (the engine's extended API).

A ``Fit`` owns the reference-layout arrays of one problem (the same arrays
``mcmcuda`` takes); a ``Session`` uploads any number of fits to the GPU once,
runs iterations with the inputs resident in HBM, and copies results back into
the fits' arrays.  A plain reference-style run is a session of one fit.

Defaults follow the reference driver ``src/imgpu.py:259-424``: mid-prior
``pcur`` (overwritten by the random start inside the prior), steps
``(pmax - pmin) / 50``, ``tcur[:, s] = 0..nbeta-1``, ``n_cnt = 1`` and the
geometric temperature ladder ``beta_i = betan ** (i / (nbeta - 1))``.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (MhreFitSpec, MhreOptions, MhreStats, ptr, MHRE_MODEL_AUTO, MHRE_MODEL_SGRA9,
                   MHRE_MODEL_SGRA13, MHRE_MODEL_USER, MHRE_MODEL_QUADSYS)

MODELS = {"auto": MHRE_MODEL_AUTO, "sgra9": MHRE_MODEL_SGRA9, "sgra13": MHRE_MODEL_SGRA13,
          "user": MHRE_MODEL_USER, "quadsys": MHRE_MODEL_QUADSYS}


def geometric_beta(nbeta, beta1=1.0, betan=1e-4):
    """Temperature ladder of src/imgpu.py:412-418 (float32)."""
    beta = np.zeros(nbeta, np.float32)
    if nbeta == 1:
        beta[0] = beta1
        return beta
    bstp = (np.float32(betan) / np.float32(beta1)) ** (1.0 / np.float32(nbeta - 1))
    for i in range(nbeta):
        beta[i] = beta1 * bstp ** i
    return beta


def var_maps(pdescr):
    """ivar / invar of src/imgpu.py:320-339."""
    pdescr = np.asarray(pdescr, np.int32)
    ivar = np.nonzero(pdescr)[0].astype(np.int32)
    invar = np.full(pdescr.size, -1, np.int32)
    invar[ivar] = np.arange(ivar.size, dtype=np.int32)
    return ivar, invar


class Fit:
    """Arrays of one MH-RE problem in the reference layouts."""

    def __init__(self, pdescr, pmint, pmaxt, coor, dat, std2r, nbeta=32, nseq=14, nburn=300, niter=500,
                 model="auto", imodel=0, icoor=None, idat=None, ptotal=None, beta=None, beta1=1.0,
                 betan=1e-4, nadj=100, npass=1, seed=0, adapt_target=0.30, pstp=None,
                 init_chains=True, model_data=False, ndatm=None, it0=0, record=True):
        f32, i32 = np.float32, np.int32
        self.model = MODELS[model] if isinstance(model, str) else int(model)
        self.imodel = int(imodel)
        self.pdescr = np.ascontiguousarray(pdescr, i32)
        self.nptot = self.pdescr.size
        self.ivar, self.invar = var_maps(self.pdescr)
        self.nprm = self.ivar.size
        self.pmint = np.ascontiguousarray(pmint, f32)
        self.pmaxt = np.ascontiguousarray(pmaxt, f32)
        self.ptotal = np.zeros(self.nptot, f32) if ptotal is None else np.ascontiguousarray(ptotal, f32)
        self.coor = np.ascontiguousarray(coor, f32)
        self.dat = np.ascontiguousarray(dat, f32)
        self.std2r = np.ascontiguousarray(std2r, f32)
        self.icoor = np.zeros(1, i32) if icoor is None else np.ascontiguousarray(icoor, i32)
        self.idat = np.zeros(1, i32) if idat is None else np.ascontiguousarray(idat, i32)
        self.nbeta, self.nseq, self.nburn, self.niter = int(nbeta), int(nseq), int(nburn), int(niter)
        self.nadj, self.npass = int(nadj), int(npass)
        self.seed = int(np.uint64(seed))
        self.adapt_target = float(adapt_target)
        self.init_chains = bool(init_chains)
        self.it0 = int(it0)
        B, S, P, N, D = self.nbeta, self.nseq, self.nprm, self.nptot, self.dat.size
        self.ndat = D
        self.beta = geometric_beta(B, beta1, betan) if beta is None else np.ascontiguousarray(beta, f32)
        pmin, pmax = self.pmint[self.ivar], self.pmaxt[self.ivar]
        self.pcur = np.tile(self.ptotal, (B, S, 1)).astype(f32)
        self.pcur[:, :, self.ivar] = (pmin + (pmax - pmin) / f32(2.0)).astype(f32)
        if pstp is None:
            step = ((pmax - pmin) / 50.0).astype(f32)
        else:
            step = np.broadcast_to(np.asarray(pstp, f32), (P,)).astype(f32)
        self.pstp = np.ascontiguousarray(np.broadcast_to(step[:, None, None], (P, B, S)), f32)
        self.tcur = np.ascontiguousarray(np.tile(np.arange(B, dtype=i32)[:, None], (1, S)))
        self.n_cnt = np.ones((P, B, S), i32)
        self.n_acpt = np.zeros((P, B, S), i32)
        self.n_exch = np.zeros((B, S), i32)
        self.n_hist = np.zeros((self.nadj, P, B, S), i32)
        self.flag = np.zeros((B, S), i32)
        self.chi2c = np.zeros((B, S), f32)
        self.ptent = np.zeros((B, S), f32)
        self.ptentn = np.zeros((B, S), i32)
        if record:
            self.pout = np.zeros((P, B, S, self.niter), f32)
            self.tout = np.zeros((B, S, self.niter), i32)
            self.chi2 = np.zeros((S, self.niter), f32)
        else:   # sampling-stage states not recorded; only the final state comes back
            self.pout = self.tout = self.chi2 = None
        self.rndst = np.zeros((B, S, 48), np.uint8)
        if model_data:
            nvis = int(self.idat[0]) if self.idat.size else 0
            self.ndatm = int(ndatm) if ndatm is not None else B * S * (D + nvis)
            self.datm = np.zeros(self.ndatm, f32)
            self.chi2m = np.zeros((B, S, D), f32)
        else:
            self.ndatm = 0
            self.datm = None
            self.chi2m = None

    def spec(self):
        F, I = C.c_float, C.c_int
        s = MhreFitSpec()
        s.model, s.imodel = self.model, self.imodel
        s.coor, s.dat, s.std2r = ptr(self.coor, F), ptr(self.dat, F), ptr(self.std2r, F)
        s.icoor, s.idat = ptr(self.icoor, I), ptr(self.idat, I)
        s.datm, s.chi2m = ptr(self.datm, F), ptr(self.chi2m, F)
        s.pdescr, s.ivar, s.invar = ptr(self.pdescr, I), ptr(self.ivar, I), ptr(self.invar, I)
        s.ptotal, s.pmint, s.pmaxt = ptr(self.ptotal, F), ptr(self.pmint, F), ptr(self.pmaxt, F)
        s.pout, s.tout, s.chi2 = ptr(self.pout, F), ptr(self.tout, I), ptr(self.chi2, F)
        s.chi2c, s.tcur, s.flag = ptr(self.chi2c, F), ptr(self.tcur, I), ptr(self.flag, I)
        s.n_acpt, s.n_exch, s.n_cnt, s.n_hist = (ptr(self.n_acpt, I), ptr(self.n_exch, I),
                                                 ptr(self.n_cnt, I), ptr(self.n_hist, I))
        s.beta, s.pstp, s.ptent, s.ptentn = ptr(self.beta, F), ptr(self.pstp, F), ptr(self.ptent, F), ptr(self.ptentn, I)
        s.pcur, s.rndst = ptr(self.pcur, F), ptr(self.rndst, C.c_ubyte)
        s.seed = self.seed
        s.ncoor, s.ndat, s.nicoor, s.nidat, s.ndatm = (self.coor.size, self.ndat, self.icoor.size,
                                                       self.idat.size, self.ndatm)
        s.nptot, s.nprm, s.nadj, s.npass = self.nptot, self.nprm, self.nadj, self.npass
        s.nbeta, s.nseq, s.nburn, s.niter = self.nbeta, self.nseq, self.nburn, self.niter
        s.adapt_target, s.init_chains, s.it0 = self.adapt_target, int(self.init_chains), self.it0
        return s


class Session:
    """Device-resident batch of fits.

    options: waves_per_block (1/2/4), persistent (0/1), graph_iters (>0 to
    replay a captured cycle), trace_iters (record decisions: ``trace`` is a dict
    of arrays for one fit, a list of such dicts for several),
    force_generic (run imodel 1/2 through the compiled-in model.cuh),
    waves_per_simd (cap on the native step kernel's resident waves per SIMD),
    streams (iterate groups of sequences on this many overlapping streams;
    0 = two in the grid form, one in the persistent form),
    exchange_launch (1: the ring+Gaussian fits exchange in a launch of their
    own after every step launch instead of inside the next one).
    """

    def __init__(self, fits, lib=None, waves_per_block=0, persistent=-1, graph_iters=0, trace_iters=0,
                 force_generic=False, waves_per_simd=0, streams=0, exchange_launch=0):
        self.L = _lib.load(lib)
        self.fits = list(fits)
        self._specs = (MhreFitSpec * len(self.fits))(*[f.spec() for f in self.fits])
        self.opt = MhreOptions()
        self.opt.waves_per_block = int(waves_per_block)
        self.opt.persistent = int(persistent)
        self.opt.graph_iters = int(graph_iters)
        self.opt.force_generic = int(bool(force_generic))
        self.opt.waves_per_simd = int(waves_per_simd)
        self.opt.streams = int(streams)
        self.opt.exchange_launch = int(bool(exchange_launch))
        self.trace = None      # single fit: dict of arrays; several fits: one dict per fit
        if trace_iters:
            T = int(trace_iters)
            nd = [T * f.nprm * f.nbeta * f.nseq for f in self.fits]
            nx = [T * f.nbeta * f.nseq for f in self.fits]
            self._tr_dec = np.full(sum(nd), -1, np.int32)
            self._tr_chi2 = np.zeros(sum(nd), np.float32)
            self._tr_ex = np.full(sum(nx), -1, np.int32)
            views, a, b = [], 0, 0
            for f, n1, n2 in zip(self.fits, nd, nx):
                views.append(dict(dec=self._tr_dec[a:a + n1].reshape(T, f.nprm, f.nbeta * f.nseq),
                                  chi2=self._tr_chi2[a:a + n1].reshape(T, f.nprm, f.nbeta * f.nseq),
                                  ex=self._tr_ex[b:b + n2].reshape(T, f.nbeta, f.nseq)))
                a += n1
                b += n2
            self.trace = views[0] if len(self.fits) == 1 else views
            self.opt.trace_iters = T
            self.opt.trace_dec = ptr(self._tr_dec, C.c_int)
            self.opt.trace_chi2 = ptr(self._tr_chi2, C.c_float)
            self.opt.trace_ex = ptr(self._tr_ex, C.c_int)
        self.h = C.c_void_p()
        _lib.check(self.L, self.L.mhre_session_create(self._specs, len(self.fits), C.byref(self.opt),
                                                      C.byref(self.h)), "mhre_session_create")

    def run(self, niter):
        _lib.check(self.L, self.L.mhre_session_run(self.h, int(niter)), "mhre_session_run")

    def time_kernels(self, niter):
        """Run niter iterations with HIP events around every launch; returns
        (step_ms, exchange_ms): one sequence group's step and exchange launch
        durations per iteration (groups overlap, so they do not add up to the
        iteration time, which stats()["run_ms"] gives after run())."""
        a, b = C.c_double(), C.c_double()
        _lib.check(self.L, self.L.mhre_session_time_kernels(self.h, int(niter), C.byref(a), C.byref(b)),
                   "mhre_session_time_kernels")
        return a.value, b.value

    def autotune(self, iters=8):
        """Pick the step kernels' launch form by timing 6 candidates for `iters`
        real iterations each; returns (waves_per_block, persistent, step_ms)."""
        w, p, t = C.c_int(), C.c_int(), C.c_double()
        _lib.check(self.L, self.L.mhre_session_autotune(self.h, int(iters), C.byref(w), C.byref(p), C.byref(t)),
                   "mhre_session_autotune")
        return w.value, p.value, t.value

    def sync(self):
        _lib.check(self.L, self.L.mhre_session_sync(self.h), "mhre_session_sync")

    def download(self, model_data=False):
        _lib.check(self.L, self.L.mhre_session_download(self.h, int(bool(model_data))), "mhre_session_download")

    def stats(self):
        st = MhreStats()
        _lib.check(self.L, self.L.mhre_session_stats(self.h, C.byref(st)), "mhre_session_stats")
        return {k: getattr(st, k) for k, _ in MhreStats._fields_}

    def fit_info(self, i=0):
        """How fit i is run: its step kernel ("sgra9", "sgra13", "user",
        "quad"), whether its replica exchange runs inside a step launch,
        whether its shift phases skip the reduction before the hardware sine,
        the session's sequence groups, and whether its quadratic-system steps
        update the residuals from row tables (one coordinate per proposal)."""
        info = (C.c_int * 5)()
        _lib.check(self.L, self.L.mhre_session_fit_info(self.h, int(i), info, 5), "mhre_session_fit_info")
        return dict(kernel=("sgra9", "sgra13", "user", "quad")[info[0]], exchange_in_step=bool(info[1]),
                    hw_phase=bool(info[2]), groups=int(info[3]), quad_rows=bool(info[4]))

    def close(self):
        if self.h:
            self.L.mhre_session_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_fits(fits, **options):
    """Run every fit to completion (nburn + niter) in one session and download."""
    with Session(fits, **options) as s:
        total = max(f.nburn + f.niter - f.it0 for f in fits)
        s.run(total)
        s.download(model_data=any(f.datm is not None for f in fits))
        return s.stats()
