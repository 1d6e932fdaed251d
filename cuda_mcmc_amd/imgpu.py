"""High-level MH-RE driver with the reference's Python API (``src/imgpu.py``).

``Mcgpu`` -- generic fits (any model compiled into the loaded library);
``Mcgpu_Sgra`` -- ring+Gaussian fits of EHT uv data.  Constructor keywords,
attributes read by scripts (``chi2, pout, chi2_2d, pout_4d, rate_acpt,
rate_exch, nvis, ncph, cphs, cpix, amp, eamp, ecph, std2r, base, phase,
ulam, vlam, ivar, invar, nprm, nptot, ptotal, nseq, niter, use_cphs, ...``)
and method return shapes follow ``src/imgpu.py:240-1067``.

Documented differences from the reference:
* ``rate_acpt`` / ``rate_exch`` sum over the sequence axis
  (``n_acpt.sum(axis=2)``, ``n_exch.sum(axis=1)``); the reference applies
  Python's builtin ``sum(x, 2)`` (sum over axis 0 plus 2, ``imgpu.py:454``).
* The device is not reset after every call (the reference calls
  ``cudaDeviceReset``, ``imgpu.py:448``): each engine call already releases
  its device memory, and a reset would also destroy state of other GPU users
  of the process.  ``reset_gpu()`` remains available.
* ``calcmodel`` passes (u, v) interleaved, the layout the model reads (the
  reference packs ``[u..., v...]``, ``imgpu.py:1048``, against
  ``model.cuh:46-48``).
* Errors raise ``EngineError`` / ``ValueError`` where the reference printed
  and returned; ``Mcgpu_Sgra`` keeps the ``error`` string attribute for its
  data checks.
"""
import time

import numpy as np

from . import mcmc_interf as mi
from .engine import geometric_beta, var_maps
from .uvdata import read_uvdata, calc_closures


class Mcgpu(object):
    """MH-RE with replica exchange for any compiled-in model (``imgpu.py:240-571``)."""

    def __init__(self, coor=None, dat=None, std2r=None, icoor=None, idat=None,
                 pdescr=None, ptotal=None, pmint=None, pmaxt=None,
                 beta=None, beta1=None, betan=None,
                 seed=None, imodel=0, npass=1,
                 nadj=100, nbeta=32, nseq=14, nburn=300, niter=500):
        f32, i32 = np.float32, np.int32
        self.dat = np.zeros(1, f32) if dat is None else np.asarray(dat, f32).ravel().copy()
        ndat = self.dat.size
        self.std2r = np.ones(ndat, f32) if std2r is None else np.asarray(std2r, f32).ravel().copy()
        self.coor = np.zeros(1, f32) if coor is None else np.asarray(coor, f32).ravel().copy()
        self.icoor = np.zeros(1, i32) if icoor is None else np.asarray(icoor, i32).ravel().copy()
        self.idat = np.zeros(1, i32) if idat is None else np.asarray(idat, i32).ravel().copy()
        self.datm = np.zeros(nbeta * nseq * ndat, f32)
        self.chi2m = np.zeros((nbeta, nseq, ndat), f32)
        if pdescr is None:
            raise ValueError("pdescr is required")
        self.pdescr = np.asarray(pdescr, i32).ravel().copy()
        self.nptot = nptot = np.int32(self.pdescr.size)
        self.ptotal = np.zeros(nptot, f32) if ptotal is None else np.asarray(ptotal, f32).ravel().copy()
        self.pmint = np.asarray(pmint, f32).ravel().copy()
        self.pmaxt = np.asarray(pmaxt, f32).ravel().copy()
        self.ivar, self.invar = var_maps(self.pdescr)
        self.nprm = nprm = np.int32(self.ivar.size)
        self.pmin = self.pmint[self.ivar]
        self.pmax = self.pmaxt[self.ivar]
        if seed is None:
            self.seed = np.uint64(np.trunc(1e6 * time.time() % (10 * nseq * nbeta)))
        else:
            self.seed = np.uint64(seed)
        self.ncoor = np.int32(self.coor.size)
        self.ndat = np.int32(ndat)
        self.nidat = np.int32(self.idat.size)
        self.nicoor = np.int32(self.icoor.size)
        self.ndatm = np.int32(self.datm.size)
        self.nadj = np.int32(nadj)
        self.nbeta = np.int32(nbeta)
        self.npass = np.int32(npass)
        self.nseq = np.int32(nseq)
        self.nburn = np.int32(nburn)
        self.niter = np.int32(niter)
        self.imodel = np.int32(imodel)
        self._alloc_state()
        self.beta1 = f32(1.0) if beta1 is None else f32(beta1)
        self.betan = f32(0.0001) if betan is None else f32(betan)
        if beta is None:
            self.beta = geometric_beta(int(nbeta), 1.0, self.betan)
        else:
            self.beta = np.asarray(beta, f32).ravel().copy()

    def _alloc_state(self):
        f32, i32 = np.float32, np.int32
        nprm, nbeta, nseq, niter = int(self.nprm), int(self.nbeta), int(self.nseq), int(self.niter)
        nadj, nptot = int(self.nadj), int(self.nptot)
        self.pstp = np.zeros((nprm, nbeta, nseq), f32)
        self.pout_4d = np.zeros((nprm, nbeta, nseq, niter), f32)
        self.tout = np.zeros((nbeta, nseq, niter), i32)
        self.chi2_2d = np.zeros((nseq, niter), f32)
        self.flag = np.zeros((nbeta, nseq), i32)
        self.tcur = np.zeros((nbeta, nseq), i32)
        self.chi2c = np.zeros((nbeta, nseq), f32)
        self.n_cnt = np.ones((nprm, nbeta, nseq), i32)
        self.n_acpt = np.zeros((nprm, nbeta, nseq), i32)
        self.n_exch = np.zeros((nbeta, nseq), i32)
        self.n_hist = np.zeros((nadj, nprm, nbeta, nseq), i32)
        self.ptentn = np.zeros((nbeta, nseq), i32)
        self.ptent = np.zeros((nbeta, nseq), f32)
        self.pcur = np.zeros((nbeta, nseq, nptot), f32)
        self.rndst = np.zeros((nbeta, nseq, 48), np.uint8)
        # prior centres (overwritten by the random start inside the prior)
        self.pcur[:, :, :] = self.ptotal
        self.pcur[:, :, self.ivar] = self.pmin + (self.pmax - self.pmin) / f32(2.0)
        self.pstp[:, :, :] = ((self.pmax - self.pmin) / 50.0).astype(f32)[:, None, None]
        self.tcur[:, :] = np.arange(nbeta, dtype=i32)[:, None]

    def burnin_and_search(self):
        mi.mcmcuda(self.coor, self.dat, self.std2r, self.icoor, self.idat, self.datm, self.chi2m,
                   self.pdescr, self.ivar, self.invar, self.ptotal, self.pmint, self.pmaxt,
                   self.pout_4d, self.tout, self.chi2_2d, self.chi2c, self.tcur, self.flag,
                   self.n_acpt, self.n_exch, self.n_cnt, self.n_hist, self.beta,
                   self.pstp, self.ptent, self.ptentn, self.pcur, self.rndst, self.seed, self.imodel,
                   self.ncoor, self.ndat, self.nicoor, self.nidat, self.ndatm, self.nptot, self.nprm,
                   self.nadj, self.npass, self.nbeta, self.nseq, self.nburn, self.niter)
        self._post_run()
        return 0

    def _post_run(self):
        nitsq = np.float32(self.niter * self.nseq)
        self.rate_acpt = self.n_acpt.sum(axis=2) / nitsq
        self.rate_exch = self.n_exch.sum(axis=1) / nitsq
        self.chi2 = self.chi2_2d.flatten()
        self.pout = self.pout_4d[:, 0, :, :].reshape((self.nprm, self.nseq * self.niter))

    def calcmodchi2(self, prm_in, imodel=None):
        """Model data and chi^2 terms for 1-D [nptot], 2-D [n, nptot] or 3-D [nbeta, nseq, nptot]
        parameter arrays (``imgpu.py:476-565``).  Returns (datm, chi2m)."""
        prm = np.array(prm_in, dtype=np.float32)
        sh = prm.shape
        if prm.ndim == 1:
            nbeta, nseq, nptot = 1, 1, sh[0]
        elif prm.ndim == 2:
            nbeta, nseq, nptot = sh[0], 1, sh[1]
        elif prm.ndim == 3:
            nbeta, nseq, nptot = sh
        else:
            raise ValueError("calcmodchi2(): input parameter array has > 3 or < 1 dimensions")
        ndat = int(self.ndat)
        pcur = np.ascontiguousarray(prm.reshape((nbeta, nseq, nptot)))
        datm = np.zeros(nbeta * nseq * ndat, np.float32)
        chi2m = np.zeros((nbeta, nseq, ndat), np.float32)
        chi2c = np.zeros((nbeta, nseq), np.float32)
        flag = np.ones((nbeta, nseq), np.int32)
        imod = self.imodel if imodel is None else imodel
        mi.calcmodchi2(self.coor, self.dat, self.std2r, self.icoor, self.idat, datm, chi2m, pcur, chi2c,
                       flag, imod, self.ncoor, self.ndat, self.nicoor, self.nidat, datm.size, nptot, nbeta, nseq)
        if len(sh) < 3:
            datm = datm.squeeze()
            chi2m = chi2m.squeeze()
        return datm, chi2m

    def reset_gpu(self):
        mi.reset_gpu()


class Mcgpu_Sgra(Mcgpu):
    """Ring+Gaussian fits of EHT uv data (``imgpu.py:585-1067``)."""

    def __init__(self, uvfile=None, imodel=1, npass=2,
                 nseq=14, nbeta=32, nburn=100, niter=100,
                 seed=1234, beta1=1., betan=0.0001,
                 pdescr=None, ptotal=None, pmint=None, pmaxt=None,
                 ulam=None, vlam=None, blin=None,
                 amp=None, phase=None, cphs=None, cpix=None, uvcp=None,
                 eamp=None, ecph=None,
                 tsec=None, chan=None, wei=None, wlam=None,
                 nadj=100, use_cphs=True):
        super(Mcgpu_Sgra, self).__init__(pdescr=pdescr, ptotal=ptotal, pmint=pmint, pmaxt=pmaxt,
                                         beta1=beta1, betan=betan, seed=seed, imodel=imodel,
                                         npass=npass, nadj=nadj, nbeta=nbeta, nseq=nseq,
                                         nburn=nburn, niter=niter)
        f32 = np.float32
        opt = lambda x, t: None if x is None else np.asarray(x, t)
        self.error = ""
        self.uvfile = uvfile
        self.imodel = np.int32(imodel)
        self.ulam, self.vlam, self.wlam = opt(ulam, f32), opt(vlam, f32), opt(wlam, f32)
        self.uvcp = opt(uvcp, f32)
        self.amp, self.phase = opt(amp, f32), opt(phase, f32)
        self.blin = opt(blin, np.int32)
        self.tsec = opt(tsec, f32)
        self.fsig = None
        self.chan = chan
        self.wei = opt(wei, f32)
        self.use_cphs = bool(use_cphs)
        if uvfile is not None:
            (self.ulam, self.vlam, self.wlam, self.amp, self.phase, self.blin, self.tsec,
             self.fsig, self.chan, self.wei) = read_uvdata(uvfile)
            self.ulam = 1e-3 * self.ulam     # Megalambda -> Gigalambda
            self.vlam = 1e-3 * self.vlam
        if self.amp is None or self.ulam is None or self.vlam is None:
            raise ValueError("uv data missing: give uvfile or ulam, vlam, amp (and phase)")
        if self.phase is None:
            self.phase = np.zeros_like(self.amp)
        if self.fsig is None:
            self.fsig = np.ones_like(self.amp)
        if np.all(self.fsig == 0):
            self.fsig[:] = 1.0
        # 1-sigma amplitude errors (imgpu.py:643-675)
        self.eamp = np.ones_like(self.amp)
        if eamp is None:
            if np.all(self.fsig != 0):
                self.eamp = np.copy(self.fsig)
            else:
                return self._fail("the uv file gives a zero sigma for some visibilities")
        elif np.isscalar(eamp):
            if eamp == 0:
                return self._fail("eamp (the amplitudes' 1-sigma error) is zero")
            self.eamp[:] = eamp
        elif len(eamp) == len(self.amp):
            if np.any(np.asarray(eamp) == 0):
                return self._fail("eamp has zero elements")
            self.eamp = np.asarray(eamp, f32).copy()
        else:
            return self._fail("eamp has %d elements for %d visibility amplitudes" % (len(eamp), len(self.amp)))
        self.base = np.sqrt(self.ulam ** 2 + self.vlam ** 2)
        self.cphs = opt(cphs, f32)
        self.cpix = opt(cpix, np.int32)
        self.trid = None
        self.tcphsec = None
        if np.all(self.phase == 0) and cphs is None:
            self.use_cphs = False
        self.nvis = nvis = np.int32(len(self.amp))
        self.nmph = nbeta * nseq * nvis
        self.cpExt = 0
        if self.use_cphs:
            if cphs is None:
                self.cpExt = 0
                if self.tsec is None or self.blin is None:
                    raise ValueError("closures need tsec and blin (or cphs with cpix/uvcp)")
                self.cphs, self.cpix, self.trid, self.tcphsec = calc_closures(self.phase, self.tsec, self.blin)
            else:
                self.cpExt = 0 if cpix is not None else 1
                if self.cpExt and self.uvcp is None:
                    raise ValueError("external closures need cpix or uvcp")
            self.ncph = ncph = np.int32(len(self.cphs))
            self.vicp = np.hstack((self.amp, self.cphs))
            self.nvicp = nvis + ncph
            self.nchi2m = nbeta * nseq * self.nvicp
            self.npass = np.int32(2)
        else:
            self.ncph = ncph = np.int32(0)
            self.vicp = np.copy(self.amp)
            self.nvicp = nvis
            self.nchi2m = nbeta * nseq * nvis
        # closure phase errors (imgpu.py:719-747)
        if self.use_cphs:
            self.ecph = np.ones(ncph, f32)
            if ecph is None:
                return self._fail("ecph (the closure phases' 1-sigma errors) is required with closure phases")
            elif np.isscalar(ecph):
                if ecph == 0:
                    return self._fail("ecph (the closure phases' 1-sigma error) is zero")
                self.ecph[:] = ecph
            elif len(ecph) == len(self.ecph):
                if np.any(np.asarray(ecph) == 0):
                    return self._fail("ecph has zero elements")
                self.ecph = np.asarray(ecph, f32).copy()
            else:
                return self._fail("ecph has %d elements for %d closure phases" % (len(ecph), len(self.cphs)))
        else:
            self.ecph = None
        # packing (imgpu.py:756-831)
        self.ndat = ndat = np.int32(nvis + ncph)
        self.nidat = np.int32(3)
        self.ndatm = np.int32(nbeta * nseq * (ndat + nvis))
        self.seed = np.uint64(seed)
        self.dat = np.empty(ndat, f32)
        self.std2r = np.ones(ndat, f32)
        self.std2r[:nvis] = 1.0 / (self.eamp.astype(np.float64) ** 2)
        if self.use_cphs:
            self.std2r[nvis:] = 1.0 / (self.ecph.astype(np.float64) ** 2)
            if self.cpExt == 0:
                self.ncoor = np.int32(2 * nvis)
                self.nicoor = np.int32(3 * ncph)
            else:
                self.ncoor = np.int32(2 * nvis + 4 * ncph)
                self.nicoor = np.int32(1)
        else:
            self.ncoor = np.int32(2 * nvis)
            self.nicoor = np.int32(1)
        self.coor = np.empty(int(self.ncoor), f32)
        self.icoor = np.zeros(int(self.nicoor), np.int32)
        self.idat = np.empty(3, np.int32)
        self.datm = np.zeros(int(self.ndatm), f32)
        self.chi2m = np.zeros((nbeta, nseq, int(ndat)), f32)
        self.mvicp = np.zeros((nbeta, nseq, int(ndat)), f32)
        self.mpha = np.zeros((nbeta, nseq, int(nvis)), f32)
        self.coor[:2 * nvis] = np.vstack((self.ulam, self.vlam)).T.flatten()
        if self.use_cphs:
            if self.cpExt:
                self.coor[2 * nvis:] = self.uvcp.flatten()
            else:
                self.icoor[:] = self.cpix.flatten()
            self.dat[:nvis] = self.amp
            self.dat[nvis:] = self.cphs
        else:
            self.dat[:] = self.amp
        self.idat[:] = nvis, self.nchi2m, self.cpExt

    def _fail(self, msg):
        self.error = msg
        return None

    def update_errors(self):
        """Recompute std2r after eamp / ecph were changed (as model_obsfit.py:283 does by hand)."""
        nvis = int(self.nvis)
        self.std2r[:nvis] = 1.0 / (self.eamp.astype(np.float64) ** 2)
        if self.use_cphs:
            self.std2r[nvis:] = 1.0 / (self.ecph.astype(np.float64) ** 2)

    def burnin_and_search(self):
        if self.error != "":
            raise ValueError("burnin_and_search: the set-up failed: " + self.error)
        Mcgpu.burnin_and_search(self)
        self.mvicp = self.datm[:self.nchi2m].reshape((self.nbeta, self.nseq, self.ndat))
        self.mpha = self.datm[self.nchi2m:].reshape((self.nbeta, self.nseq, self.nvis))
        return 0

    def calcmodchi2(self, prm_in, imodel=None):
        """Model amplitudes/closures and chi^2 for parameter arrays (``imgpu.py:907-985``).

        Returns chi2c, mvicp, mpha, chi2m, datm."""
        if self.error != "":
            raise ValueError("the set-up failed: " + self.error)
        prm = np.array(prm_in, dtype=np.float32)
        sh = prm.shape
        if len(sh) == 1:
            nbeta, nseq, nptot = 1, 1, sh[0]
        elif len(sh) == 2:
            nbeta, nseq, nptot = sh[0], 1, sh[1]
        elif len(sh) == 3:
            nbeta, nseq, nptot = sh
        else:
            raise ValueError("calcmodchi2(): input parameter array has > 3 or < 1 dimensions")
        ndat, nvis = int(self.ndat), int(self.nvis)
        pcur = np.ascontiguousarray(prm.reshape((nbeta, nseq, nptot)))
        ndatm = nbeta * nseq * (ndat + nvis)
        datm = np.zeros(ndatm, np.float32)
        chi2m = np.zeros((nbeta, nseq, ndat), np.float32)
        chi2c = np.zeros((nbeta, nseq), np.float32)
        flag = np.ones((nbeta, nseq), np.int32)
        imod = imodel if imodel in (1, 2) else self.imodel
        idat = self.idat.copy()
        idat[1] = nbeta * nseq * ndat
        mi.calcmodchi2(self.coor, self.dat, self.std2r, self.icoor, idat, datm, chi2m, pcur, chi2c, flag,
                       imod, self.ncoor, self.ndat, self.nicoor, self.nidat, ndatm, nptot, nbeta, nseq)
        nchi2m = nbeta * nseq * ndat
        mvicp = datm[:nchi2m].reshape((nbeta, nseq, ndat))
        mpha = datm[nchi2m:].reshape((nbeta, nseq, nvis))
        if len(sh) == 1:
            chi2c = chi2c[0, 0]
            mvicp = mvicp.flatten()
            mpha = mpha.flatten()
            chi2m = chi2m.flatten()
        return chi2c, mvicp, mpha, chi2m, datm

    def calcmodel(self, ulam, vlam, prm_in, imodel=2):
        """Model amplitudes and phases at (ulam, vlam) (``imgpu.py:988-1067``).

        ulam, vlam: 2-D grids; prm_in: [nptot], [n, nptot] or [nbeta, nseq, nptot].
        Returns mamp, mpha shaped like ulam (1-D prm), [n, ...] or [nbeta, nseq, ...]."""
        prm = np.array(prm_in, dtype=np.float32)
        ulam = np.asarray(ulam, np.float32)
        vlam = np.asarray(vlam, np.float32)
        ul, vl = ulam.flatten(), vlam.flatten()
        uv_sh = ulam.shape
        if prm.ndim == 1:
            nbeta, nseq, nptot = 1, 1, prm.shape[0]
            mamp_sh = uv_sh
        elif prm.ndim == 2:
            nbeta, nseq, nptot = prm.shape[0], 1, prm.shape[1]
            mamp_sh = (nbeta,) + uv_sh
        elif prm.ndim == 3:
            nbeta, nseq, nptot = prm.shape
            mamp_sh = (nbeta, nseq) + uv_sh
        else:
            raise ValueError("calcmodel(): input parameter array has > 3 or < 1 dimensions")
        imod = imodel if imodel in (1, 2) else self.imodel
        nvis = ul.size
        coor = np.empty(2 * nvis, np.float32)
        coor[0::2] = ul
        coor[1::2] = vl
        ndatm = nbeta * nseq * 2 * nvis
        datm = np.zeros(ndatm, np.float32)
        idat = np.array((nvis, 0, 0), dtype=np.int32)
        pcur = np.ascontiguousarray(prm.reshape((nbeta, nseq, nptot)))
        mi.calcmodel(coor, idat, datm, pcur, imod, coor.size, nvis, idat.size, ndatm, nptot, nbeta, nseq)
        nm = nbeta * nseq * nvis
        return datm[:nm].reshape(mamp_sh), datm[nm:].reshape(mamp_sh)
