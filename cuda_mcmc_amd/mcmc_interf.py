"""Python <-> engine interface with the reference's function names and arguments.

Drop-in for the Cython module ``src/mcmc_interf.pyx``: ``mcmcuda``,
``calcmodchi2``, ``calcmodel`` and ``reset_gpu`` take the same arguments in
the same order with the same dtypes and dimensionalities
(``src/mcmc_interf.pyx:230-343``) and fill the output arrays in place.
Differences: arguments are checked for dtype, dimensionality, contiguity and
writeability (the reference casts ``.data`` blindly), and failures raise
``EngineError`` instead of terminating the process.
"""
import numpy as np

from . import _lib
from ._lib import EngineError, need, ptr
import ctypes as C

f32, i32, u8 = np.float32, np.int32, np.uint8


def mcmcuda(coor, dat, std2r, icoor, idat, datm, chi2m,
            pdescr, ivar, invar, ptotal, pmint, pmaxt,
            pout, tout, chi2, chi2c, tcur, flag,
            n_acpt, n_exch, n_cnt, n_hist, beta,
            pstp, ptent, ptentn, pcur, rndst,
            seed, imodel, ncoor, ndat, nicoor, nidat, ndatm, nptot, nprm,
            nadj, npass, nbeta, nseq, nburn, niter):
    """MH-RE fit (src/mcmc_interf.pyx:230-281 -> mcmc_cuda)."""
    L = _lib.load()
    arrs = [
        (coor, f32, 1, "coor"), (dat, f32, 1, "dat"), (std2r, f32, 1, "std2r"),
        (icoor, i32, 1, "icoor"), (idat, i32, 1, "idat"), (datm, f32, 1, "datm"),
        (chi2m, f32, 3, "chi2m"), (pdescr, i32, 1, "pdescr"), (ivar, i32, 1, "ivar"),
        (invar, i32, 1, "invar"), (ptotal, f32, 1, "ptotal"), (pmint, f32, 1, "pmint"),
        (pmaxt, f32, 1, "pmaxt"), (pout, f32, 4, "pout"), (tout, i32, 3, "tout"),
        (chi2, f32, 2, "chi2"), (chi2c, f32, 2, "chi2c"), (tcur, i32, 2, "tcur"),
        (flag, i32, 2, "flag"), (n_acpt, i32, 3, "n_acpt"), (n_exch, i32, 2, "n_exch"),
        (n_cnt, i32, 3, "n_cnt"), (n_hist, i32, 4, "n_hist"), (beta, f32, 1, "beta"),
        (pstp, f32, 3, "pstp"), (ptent, f32, 2, "ptent"), (ptentn, i32, 2, "ptentn"),
        (pcur, f32, 3, "pcur"), (rndst, u8, 3, "rndst"),
    ]
    for a, dt, nd, name in arrs:
        need(a, dt, nd, name)
    B, S, P, N = int(nbeta), int(nseq), int(nprm), int(nptot)
    shapes = dict(pout=(P, B, S, int(niter)), tout=(B, S, int(niter)), chi2=(S, int(niter)),
                  chi2c=(B, S), tcur=(B, S), flag=(B, S), n_acpt=(P, B, S), n_exch=(B, S),
                  n_cnt=(P, B, S), n_hist=(int(nadj), P, B, S), pstp=(P, B, S), ptent=(B, S),
                  ptentn=(B, S), pcur=(B, S, N), rndst=(B, S, 48), chi2m=(B, S, int(ndat)))
    loc = locals()
    for name, shp in shapes.items():
        if loc[name].shape != shp:
            raise ValueError("%s has shape %s, expected %s" % (name, loc[name].shape, shp))
    sizes = dict(coor=ncoor, dat=ndat, std2r=ndat, icoor=nicoor, idat=nidat, datm=ndatm,
                 pdescr=nptot, invar=nptot, ptotal=nptot, pmint=nptot, pmaxt=nptot, ivar=nprm, beta=nbeta)
    for name, n in sizes.items():
        if loc[name].size < int(n):
            raise ValueError("%s has %d elements, %s requires %d" % (name, loc[name].size, name, int(n)))
    F, I = C.c_float, C.c_int
    rc = L.mcmc_cuda(ptr(coor, F), ptr(dat, F), ptr(std2r, F), ptr(icoor, I), ptr(idat, I),
                     ptr(datm, F), ptr(chi2m, F), ptr(pdescr, I), ptr(ivar, I), ptr(invar, I),
                     ptr(ptotal, F), ptr(pmint, F), ptr(pmaxt, F), ptr(pout, F), ptr(tout, I),
                     ptr(chi2, F), ptr(chi2c, F), ptr(tcur, I), ptr(flag, I), ptr(n_acpt, I),
                     ptr(n_exch, I), ptr(n_cnt, I), ptr(n_hist, I), ptr(beta, F), ptr(pstp, F),
                     ptr(ptent, F), ptr(ptentn, I), ptr(pcur, F), ptr(rndst, C.c_ubyte),
                     int(np.uint64(seed)), int(imodel), int(ncoor), int(ndat), int(nicoor), int(nidat),
                     int(ndatm), int(nptot), int(nprm), int(nadj), int(npass), int(nbeta), int(nseq),
                     int(nburn), int(niter))
    _lib.check(L, rc, "mcmc_cuda")


def calcmodchi2(coor, dat, std2r, icoor, idat, datm, chi2m, pcur, chi2c, flag,
                imodel, ncoor, ndat, nicoor, nidat, ndatm, nptot, nbeta, nseq):
    """Batched model data and chi^2 (src/mcmc_interf.pyx:294-313 -> calc_modchi2)."""
    L = _lib.load()
    for a, dt, nd, name in [(coor, f32, 1, "coor"), (dat, f32, 1, "dat"), (std2r, f32, 1, "std2r"),
                            (icoor, i32, 1, "icoor"), (idat, i32, 1, "idat"), (datm, f32, 1, "datm"),
                            (chi2m, f32, 3, "chi2m"), (pcur, f32, 3, "pcur"), (chi2c, f32, 2, "chi2c"),
                            (flag, i32, 2, "flag")]:
        need(a, dt, nd, name)
    if pcur.shape != (int(nbeta), int(nseq), int(nptot)):
        raise ValueError("pcur has shape %s, expected %s" % (pcur.shape, (nbeta, nseq, nptot)))
    F, I = C.c_float, C.c_int
    rc = L.calc_modchi2(ptr(coor, F), ptr(dat, F), ptr(std2r, F), ptr(icoor, I), ptr(idat, I),
                        ptr(datm, F), ptr(chi2m, F), ptr(pcur, F), ptr(chi2c, F), ptr(flag, I),
                        int(imodel), int(ncoor), int(ndat), int(nicoor), int(nidat), int(ndatm),
                        int(nptot), int(nbeta), int(nseq))
    _lib.check(L, rc, "calc_modchi2")


def calcmodel(coor, idat, datm, pcur, imodel, ncoor, ndat, nidat, ndatm, nptot, nbeta, nseq):
    """Model data only (src/mcmc_interf.pyx:325-336 -> calc_model).  coor holds (u, v) pairs."""
    L = _lib.load()
    for a, dt, nd, name in [(coor, f32, 1, "coor"), (idat, i32, 1, "idat"), (datm, f32, 1, "datm"),
                            (pcur, f32, 3, "pcur")]:
        need(a, dt, nd, name)
    F, I = C.c_float, C.c_int
    rc = L.calc_model(ptr(coor, F), ptr(idat, I), ptr(datm, F), ptr(pcur, F), int(imodel), int(ncoor),
                      int(ndat), int(nidat), int(ndatm), int(nptot), int(nbeta), int(nseq))
    _lib.check(L, rc, "calc_model")


def reset_gpu():
    """Release the device (src/mcmc_interf.pyx:342-343)."""
    L = _lib.load()
    _lib.check(L, L.reset_gpu(), "reset_gpu")


def list_gpus():
    """Visible GPUs, one line each (replaces src/list_gpus.cu)."""
    L = _lib.load()
    buf = C.create_string_buffer(8192)
    _lib.check(L, L.list_gpus(buf, len(buf)), "list_gpus")
    return buf.value.decode()
