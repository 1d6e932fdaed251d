"""Post-processing of fits: model images and parameter-histogram summaries.

model_vis2bri_simple restates ``src/obsdatafit.py:2683-2804``: the model's
visibilities on an ngrid x ngrid uv grid (U axis right to left, as in the
reference), multiplied by the cosine window 0.5 (1 + cos(pi rho / (UVspan/2)))
(zero beyond UVspan/2), inverse-FFT'd to the brightness image, with the uv
and xy grids and extents.  As in the reference, for even ngrid the U grid
runs from UVspan/2 - du down to -UVspan/2, so the cell that the FFT takes as
zero frequency holds u = -du.  For odd ngrid the reference leaves the V
range undefined; here it is linspace(-UVspan/2, UVspan/2, ngrid).

model_vis2bri restates ``src/obsdatafit.py:2540-2680``: the same window on an
ntotal x ntotal grid (for odd ntotal the V range is the reversed U range),
``ifftshift(ifft2(.))`` without the inner fftshift of the simple form (so the
image carries the phase ramp of the grid's offset zero-frequency cell; its
modulus is the brightness), then central cuts: nb x nb of the image, nv x nv
of the visibilities and grids.  As in the reference, the X/Y grids are cut to
nv (not nb) pixels.

The grid conventions (U and X running right to left, even grids starting one
cell short of +UVspan/2, the cuts' index arithmetic) are the output format that
callers plot with and compare across runs, so both functions follow the
reference statement by statement.

param_histograms and histogram_mean_params restate the summaries that the EHT
driver draws from the coldest-slot samples
(``examples/sgra_model_fit/model_obsfit.py:288-298``, ``:412-440``,
``:515-526``): per varied parameter the 50-bin histogram, mean, standard
deviation and the best-fit value at the minimum recorded chi^2, and the
parameter vector of histogram means (fixed parameters from ptotal).
histogram_mean_image adds the model image of that vector and its chi^2 and
reduced chi^2 chi^2 / (nvis + ncph - nprm - 1) (``:528-562``).

gpu_visfun(imodel) gives a visfun evaluated on the GPU through calc_model
(the engine's 9- or 13-parameter model), for use as the first argument.
"""
import numpy as np

from . import mcmc_interf


def model_vis2bri_simple(visfun, vfparam, UVspan, ngrid):
    """Returns Vis, Bri, Ugrid, Vgrid, Xgrid, Ygrid, UVext, XYext (see module doc)."""
    uvext = UVspan / 2.0
    if ngrid % 2 == 0:
        urng = np.linspace(uvext - UVspan / ngrid, -uvext, ngrid)
        vrng = np.linspace(-uvext, uvext - UVspan / ngrid, ngrid)
    else:
        urng = np.linspace(uvext, -uvext, ngrid)
        vrng = np.linspace(-uvext, uvext, ngrid)
    Ugrid, Vgrid = np.meshgrid(urng, vrng)
    rho = np.sqrt(Ugrid ** 2 + Vgrid ** 2)
    cwin = 0.5 * (1.0 + np.cos(np.pi * rho / uvext))
    cwin[rho > uvext] = 0.0
    Vi = visfun(Ugrid, Vgrid, vfparam)
    Br = np.fft.ifftshift(np.fft.ifft2(np.fft.fftshift(Vi * cwin)))
    XYspan = 3600.0 * np.degrees(ngrid / UVspan) * 1e-3    # nanoradians -> microarcseconds
    xyext = XYspan / 2.0
    if ngrid % 2 == 0:
        xrng = np.linspace(xyext - XYspan / ngrid, -xyext, ngrid)
        yrng = np.linspace(-xyext, xyext - XYspan / ngrid, ngrid)
    else:
        xrng = np.linspace(xyext, -xyext, ngrid)
        yrng = np.linspace(-xyext, xyext, ngrid)
    UVext = (Ugrid[0, 0], Ugrid[0, -1], Vgrid[0, 0], Vgrid[-1, 0])
    Xgrid, Ygrid = np.meshgrid(xrng, yrng)
    XYext = (Xgrid[0, 0], Xgrid[0, -1], Ygrid[0, 0], Ygrid[-1, 0])
    return Vi, Br, Ugrid, Vgrid, Xgrid, Ygrid, UVext, XYext


def gpu_visfun(imodel=1):
    """visfun(U, V, prm) -> complex visibilities of one parameter vector on the GPU."""
    def visfun(U, V, prm):
        U = np.asarray(U, np.float32)
        V = np.asarray(V, np.float32)
        prm = np.ascontiguousarray(prm, np.float32).reshape(1, 1, -1)
        n = U.size
        coor = np.empty(2 * n, np.float32)
        coor[0::2] = U.ravel()
        coor[1::2] = V.ravel()
        idat = np.array((n, 0, 0), np.int32)
        datm = np.zeros(2 * n, np.float32)
        mcmc_interf.calcmodel(coor, idat, datm, prm, int(imodel), coor.size, n, idat.size, datm.size,
                              prm.shape[-1], 1, 1)
        amp = datm[:n].astype(np.float64).reshape(U.shape)
        pha = datm[n:].astype(np.float64).reshape(U.shape)
        return amp * np.exp(1j * pha)
    return visfun


def model_vis2bri(visfun, vfparam, UVspan_total, ntotal, nv, nb):
    """Returns Vis (nv x nv cut), Bri (nb x nb cut), Ugrid, Vgrid, Xgrid, Ygrid,
    UVext, XYext (see module doc; src/obsdatafit.py:2540-2680)."""
    uvext = UVspan_total / 2.0
    if ntotal % 2 == 0:
        urng = np.linspace(uvext - UVspan_total / ntotal, -uvext, ntotal)
        vrng = np.linspace(-uvext, uvext - UVspan_total / ntotal, ntotal)
    else:
        urng = np.linspace(uvext, -uvext, ntotal)
        vrng = urng[::-1]
    Ugrid_total, Vgrid_total = np.meshgrid(urng, vrng)
    rho = np.sqrt(Ugrid_total ** 2 + Vgrid_total ** 2)
    cwin = 0.5 * (1.0 + np.cos(np.pi * rho / uvext))
    cwin[rho > uvext] = 0.0
    Vi = np.asarray(visfun(Ugrid_total.ravel(), Vgrid_total.ravel(), vfparam)).reshape(ntotal, ntotal)
    Br = np.fft.ifftshift(np.fft.ifft2(Vi * cwin))
    XYspan = 3600.0 * np.degrees(ntotal / UVspan_total) * 1e-3    # nanoradians -> microarcseconds
    xyext = XYspan / 2.0
    if ntotal % 2 == 0:
        xrng_t = np.linspace(xyext - XYspan / ntotal, -xyext, ntotal)
        yrng_t = np.linspace(-xyext, xyext - XYspan / ntotal, ntotal)
    else:
        yrng_t = np.linspace(-xyext, xyext, ntotal)
        xrng_t = yrng_t[::-1]
    m = ntotal // 2

    def cut(n):
        k = n // 2
        return slice(m - k, m + k + (n % 2))

    Bri = Br[cut(nb), cut(nb)]
    sv = cut(nv)
    Ugrid = Ugrid_total[sv, sv]
    Vgrid = Vgrid_total[sv, sv]
    Vis = Vi[sv, sv]
    UVext = (Ugrid[0, 0], Ugrid[0, -1], Vgrid[0, 0], Vgrid[-1, 0])
    Xgrid, Ygrid = np.meshgrid(xrng_t[sv], yrng_t[sv])
    XYext = (Xgrid[0, 0], Xgrid[0, -1], Ygrid[0, 0], Ygrid[-1, 0])
    return Vis, Bri, Ugrid, Vgrid, Xgrid, Ygrid, UVext, XYext


def coldest_samples(pout_4d, chi2_2d):
    """pout1 [nprm, nseq*niter] of the beta = 1 slot and the index of the
    minimum recorded chi^2 (model_obsfit.py:288-298)."""
    pout_4d = np.asarray(pout_4d)
    nprm = pout_4d.shape[0]
    pout1 = pout_4d[:, 0, :, :].reshape(nprm, -1).astype(np.float32)
    im = int(np.asarray(chi2_2d).ravel().argmin())
    return pout1, im


def param_histograms(pout_4d, chi2_2d, ivar, pdescr, nbins=50):
    """Per varied parameter i: dict(counts, edges, mean, std, bestfit) of the
    coldest-slot samples, angles (pdescr == 2) in degrees
    (model_obsfit.py:412-440).  The reference tests pdescr[i] with the varied
    index i; here the parameter's own entry pdescr[ivar[i]] is used, which is
    the same when every parameter is varied."""
    pout1, im = coldest_samples(pout_4d, chi2_2d)
    out = []
    for i in range(pout1.shape[0]):
        pp = pout1[i, :].astype(np.float64)
        if int(pdescr[int(ivar[i])]) == 2:
            pp = np.degrees(pp)
        counts, edges = np.histogram(pp, nbins)
        out.append(dict(counts=counts, edges=edges, mean=float(np.mean(pp)), std=float(np.std(pp)),
                        bestfit=float(pp[im])))
    return out


def histogram_mean_params(pout_4d, chi2_2d, ptotal, invar):
    """The parameter vector of histogram means (model_obsfit.py:515-526):
    varied parameters take the mean of their coldest-slot samples (angles
    averaged as numbers, as in the reference), fixed ones their ptotal value."""
    pout1, _ = coldest_samples(pout_4d, chi2_2d)
    mus = pout1.mean(axis=1)
    prm = np.array(ptotal, dtype=np.float64).copy()
    for i, iv in enumerate(np.asarray(invar)):
        if iv >= 0:
            prm[i] = mus[iv]
    return prm


def histogram_mean_image(sgra, UVspan, ngrid, visfun=None):
    """Model image of the histogram-mean parameters of a finished
    imgpu.Mcgpu_Sgra run, with its chi^2 and reduced chi^2
    (model_obsfit.py:515-562).  Returns dict(prm, Vi, Br, Ugrid, Vgrid, Xgrid,
    Ygrid, UVext, XYext, chi2, chi2_reduced)."""
    prm = histogram_mean_params(sgra.pout_4d, sgra.chi2_2d, sgra.ptotal, sgra.invar)
    vf = visfun or gpu_visfun(int(sgra.imodel))
    Vi, Br, Ug, Vg, Xg, Yg, UVext, XYext = model_vis2bri_simple(vf, prm, UVspan, ngrid)
    chi2 = float(np.ravel(sgra.calcmodchi2(prm.astype(np.float32))[0])[0])
    dof = sgra.nvis + sgra.ncph - sgra.nprm - 1
    return dict(prm=prm, Vi=Vi, Br=Br, Ugrid=Ug, Vgrid=Vg, Xgrid=Xg, Ygrid=Yg, UVext=UVext, XYext=XYext,
                chi2=chi2, chi2_reduced=chi2 / dof)
