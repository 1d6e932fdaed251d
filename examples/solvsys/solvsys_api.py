"""The solvsys problem through the raw mcmc_interf.mcmcuda call, building every
array by hand (Python 3 counterpart of the reference's
examples/solvsys/solvsys_api.py).
"""
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
os.environ.setdefault("MHRE_LIB", os.path.join(HERE, "libmhre.so"))

import numpy as np  # noqa: E402
from numpy import array, float32, int32, uint8, uint64, zeros, ones_like, arange  # noqa: E402

from cuda_mcmc_amd import mcmc_interf as mi  # noqa: E402

nseq, nbeta, niter, nburn = int32(14), int32(32), int32(5000), int32(400)
seed = uint64(np.trunc(1e6 * time.time() % (10 * nseq * nbeta)))
imodel = 0
pdescr = array((1, 1), dtype=int32)
ptotal = array((0., 0.), dtype=float32)
nptot = len(ptotal)
pmint = -5 * ones_like(ptotal)
pmaxt = 5 * ones_like(ptotal)
ivar = np.where(pdescr != 0)[0].astype(int32)
nprm = len(ivar)
pmin, pmax = pmint[ivar], pmaxt[ivar]
invar = np.full(nptot, -1, int32)
invar[ivar] = arange(nprm, dtype=int32)
nadj = int32(100)
betan = float32(0.0001)
npass = 1

coor = array((2, 2, 0.25, 0.5), dtype=float32)
icoor = zeros(1, dtype=int32)
dat = array((1, 0.5), dtype=float32)
std2r = ones_like(dat, dtype=float32)
idat = zeros(1, dtype=int32)
ndat = len(dat)
datm = zeros(nbeta * nseq * ndat, dtype=float32)
chi2m = zeros((nbeta, nseq, ndat), dtype=float32)
rndst = zeros((nbeta, nseq, 48), dtype=uint8)
beta = zeros(nbeta, dtype=float32)
pstp = zeros((nprm, nbeta, nseq), dtype=float32)
pout = zeros((nprm, nbeta, nseq, niter), dtype=float32)
tout = zeros((nbeta, nseq, niter), dtype=int32)
chi2 = zeros((nseq, niter), dtype=float32)
flag = zeros((nbeta, nseq), dtype=int32)
tcur = zeros((nbeta, nseq), dtype=int32)
chi2c = zeros((nbeta, nseq), dtype=float32)
n_cnt = np.ones((nprm, nbeta, nseq), dtype=int32)
n_acpt = zeros((nprm, nbeta, nseq), dtype=int32)
n_exch = zeros((nbeta, nseq), dtype=int32)
n_hist = zeros((nadj, nprm, nbeta, nseq), dtype=int32)
ptentn = zeros((nbeta, nseq), dtype=int32)
ptent = zeros((nbeta, nseq), dtype=float32)
pcur = zeros((nbeta, nseq, nptot), dtype=float32)
pcur[:, :, :] = ptotal
pcur[:, :, ivar] = pmin + (pmax - pmin) / 2.
pstp[:, :, :] = ((pmax - pmin) / 50.)[:, None, None]
tcur[:, :] = arange(nbeta)[:, None]
bstp = (betan / 1.0) ** (1.0 / float(nbeta - 1))
beta[:] = (bstp ** arange(nbeta)).astype(float32)

mi.mcmcuda(coor, dat, std2r, icoor, idat, datm, chi2m, pdescr, ivar, invar, ptotal, pmint, pmaxt,
           pout, tout, chi2, chi2c, tcur, flag, n_acpt, n_exch, n_cnt, n_hist, beta, pstp, ptent, ptentn,
           pcur, rndst, seed, imodel, len(coor), ndat, len(icoor), len(idat), len(datm), nptot, nprm,
           nadj, npass, nbeta, nseq, nburn, niter)

c = chi2.flatten()
p = pout[:, 0, :, :].reshape(nprm, nseq * niter)
im = c.argmin()
print("The minimum chi^2 at chi2[%d] = %f, (x, y) = (%f, %f)" % (im, c[im], p[0, im], p[1, im]))
print("acceptance rates (beta = 1):", n_acpt[:, 0, :].sum(axis=1) / float(niter * nseq))
