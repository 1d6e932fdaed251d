"""Model visibilities to a brightness image (post-processing of fits).

model_vis2bri_simple restates ``src/obsdatafit.py:2683-2804``: the model's
visibilities on an ngrid x ngrid uv grid (U axis right to left, as in the
reference), multiplied by the cosine window 0.5 (1 + cos(pi rho / (UVspan/2)))
(zero beyond UVspan/2), inverse-FFT'd to the brightness image, with the uv
and xy grids and extents.  As in the reference, for even ngrid the U grid
runs from UVspan/2 - du down to -UVspan/2, so the cell that the FFT takes as
zero frequency holds u = -du.  For odd ngrid the reference leaves the V
range undefined; here it is linspace(-UVspan/2, UVspan/2, ngrid).

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
