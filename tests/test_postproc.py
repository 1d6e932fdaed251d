"""Visibility -> brightness post-processing (model_vis2bri_simple).

CPU: with the oracle's float64 model as visfun, the image sums to the
windowed visibility of the grid cell that the FFT treats as zero frequency
(only that term survives the sum over pixels; with the reference's
right-to-left U grid that cell is u = -du, v = 0), the uv and xy grids are
each other's Fourier duals, and the image of an unrotated source is
mirror-symmetric in y.
"""
import numpy as np

from cuda_mcmc_amd.postproc import model_vis2bri_simple, model_vis2bri
from oracle.models import sgra_model9

PRM = np.array([2.0, 25.0, 0.6, 0.3, 0.4, 3.0, 0.8, 0.2, 0.0])


def _vf(U, V, p):
    return sgra_model9(U.ravel(), V.ravel(), p).reshape(U.shape)


def test_image_flux_and_grids():
    Vi, Br, Ug, Vg, Xg, Yg, UVext, XYext = model_vis2bri_simple(_vf, PRM, 20.0, 128)
    assert Vi.shape == Br.shape == (128, 128)
    rho = np.sqrt(Ug ** 2 + Vg ** 2)
    cwin = np.where(rho > 10.0, 0.0, 0.5 * (1.0 + np.cos(np.pi * rho / 10.0)))
    np.testing.assert_allclose(Br.sum(), (Vi * cwin)[64, 64], rtol=1e-9)
    assert abs(Vg[64, 64]) < 1e-12 and abs(Ug[64, 64] + 20.0 / 128) < 1e-12
    np.testing.assert_allclose(abs((Vi * cwin)[64, 64]), PRM[0], rtol=1e-2)   # a short baseline: ~Zsp
    # pixel sizes: du * dx = 1 / ngrid (in radians x lambda)
    du = (Ug[0, 0] - Ug[0, 1]) * 1e9
    dx = (Xg[0, 0] - Xg[0, 1]) * np.pi / 180 / 3600 * 1e-6
    np.testing.assert_allclose(du * dx * 128, 1.0, rtol=1e-12)
    assert UVext[0] > UVext[1] and XYext[0] > XYext[1]     # U and X run right to left


def test_unrotated_source_is_symmetric_in_y():
    Vi, Br, *_ = model_vis2bri_simple(_vf, PRM, 20.0, 128)
    img = Br.real
    # th = 0: the brightness is even in y (about the centre row, index 64)
    np.testing.assert_allclose(img[65:, :], img[63:0:-1, :], atol=1e-9 * np.abs(img).max())


# ------------------------------------------------ images against the analytic source

def _gaussian_image(X, Y, p):
    """Brightness (flux per nrad^2) of the 9-parameter model with gq = 1 and
    the Gaussian centred (Ri = d): an elliptical Gaussian of FWHM gsx along
    x' = x cos th + y sin th and gsy along y', total flux Zsp."""
    from oracle.models import UAS2NRAD
    Zsp, Re, rq, ecc, fade, gax, aq, gq, th = p
    Rex = Re * UAS2NRAD
    gsy = gax * Rex
    gsx = aq * gsy
    x, y = X * UAS2NRAD, Y * UAS2NRAD
    xp = x * np.cos(th) + y * np.sin(th)
    yp = -x * np.sin(th) + y * np.cos(th)
    k = 2.0 * np.sqrt(2.0 * np.log(2.0))
    sx, sy = gsx / k, gsy / k
    return Zsp / (2 * np.pi * sx * sy) * np.exp(-0.5 * (xp / sx) ** 2 - 0.5 * (yp / sy) ** 2)


ECC = 0.25
GAUSS = np.array([1.5, 25.0, ECC / (1 + ECC), ECC, 0.4, 2.5, 0.6, 1.0, 0.5])   # Ri = d: centred


def _vflat(U, V, p):
    return sgra_model9(np.ravel(U), np.ravel(V), p).reshape(np.shape(U))


def test_simple_image_is_the_analytic_gaussian():
    """|Br| pixel values are the source's flux per pixel at the grid coordinates.

    The uv span (200 Glambda) makes the cosine window ~1 where the Gaussian's
    visibility lives (0.1 % effect).  The grid's zero-frequency cell is
    u = -du (U runs right to left from +UVspan/2 - du), so pixel (X, Y) holds
    the brightness at (X + dx, Y); the mirrored source (-X) does not fit."""
    from oracle.models import UAS2NRAD
    Vi, Br, Ug, Vg, Xg, Yg, UVext, XYext = model_vis2bri_simple(_vflat, GAUSS, 200.0, 256)
    dxu = Xg[0, 0] - Xg[0, 1]
    dx = dxu * UAS2NRAD
    A = _gaussian_image(Xg + dxu, Yg, GAUSS) * dx * dx
    assert np.abs(np.abs(Br) - A).max() < 3e-3 * A.max()
    M = _gaussian_image(-Xg, Yg, GAUSS) * dx * dx
    assert np.abs(np.abs(Br) - M).max() > 0.2 * A.max()
    np.testing.assert_allclose(np.abs(Br).sum(), GAUSS[0], rtol=1e-6)


def test_central_cut_image():
    """model_vis2bri (obsdatafit.py:2540-2680): the nb x nb central cut of the
    same image (its modulus; the complex values carry the phase ramp of the
    grid's offset origin) and the nv x nv cuts of visibilities and grids."""
    from oracle.models import UAS2NRAD
    Vis, Bri, Ug, Vg, Xg, Yg, UVext, XYext = model_vis2bri(_vflat, GAUSS, 200.0, 256, 96, 80)
    assert Bri.shape == (80, 80) and Vis.shape == Ug.shape == Xg.shape == (96, 96)
    # the X/Y grids are cut to nv like the visibilities (as in the reference)
    Vis2, Bri2, Ug2, Vg2, Xg2, Yg2, _, _ = model_vis2bri(_vflat, GAUSS, 200.0, 256, 80, 80)
    dxu = Xg2[0, 0] - Xg2[0, 1]
    dx = dxu * UAS2NRAD
    A = _gaussian_image(Xg2 + dxu, Yg2, GAUSS) * dx * dx
    assert np.abs(np.abs(Bri2) - A).max() < 3e-3 * A.max()
    np.testing.assert_allclose(np.abs(Bri2).sum(), A.sum(), rtol=3e-3)
    # the visibility cut is the model on the cut grid
    np.testing.assert_allclose(Vis2, _vflat(Ug2, Vg2, GAUSS), rtol=1e-12)
    assert abs(Vg2[40, 40]) < 1e-12 and abs(Ug2[40, 40] + 200.0 / 256) < 1e-12
    assert UVext[0] > UVext[1] and XYext[0] > XYext[1]
    # odd sizes: the V range is the reversed U range and the cut is centred
    Vis3, Bri3, Ug3, Vg3, *_ = model_vis2bri(_vflat, GAUSS, 200.0, 255, 33, 33)
    assert Bri3.shape == (33, 33) and abs(Ug3[16, 16]) < 1e-12 and abs(Vg3[16, 16]) < 1e-12


# ------------------------------------------------------- parameter histograms

def test_histogram_summaries_known_answers():
    """Samples of the coldest slot pout_4d[:, 0] with known moments: parameter 0
    alternates 1, 3 (mean 2, std 1), parameter 1 is an angle (pdescr 2) equal
    to pi/2 (90 degrees), and the best fit is read at the minimum chi^2."""
    from cuda_mcmc_amd.postproc import param_histograms, histogram_mean_params
    nprm, nbeta, nseq, niter = 2, 3, 4, 10
    pout = np.zeros((nprm, nbeta, nseq, niter), np.float32)
    pout[0, 0] = np.where((np.arange(nseq * niter) % 2).reshape(nseq, niter) == 0, 1.0, 3.0)
    pout[1, 0] = np.pi / 2
    pout[:, 1:] = 99.0                       # hotter slots are not part of the summary
    chi2 = np.full((nseq, niter), 10.0, np.float32)
    chi2[2, 7] = 1.0                         # sample 2 * niter + 7 = 27 (odd: parameter 0 = 3)
    pdescr = np.array([1, 0, 2], np.int32)   # parameter 1 of nptot 3 is fixed
    ivar = np.array([0, 2], np.int32)
    invar = np.array([0, -1, 1], np.int32)
    h = param_histograms(pout, chi2, ivar, pdescr, nbins=50)
    assert h[0]["mean"] == 2.0 and h[0]["std"] == 1.0 and h[0]["bestfit"] == 3.0
    assert h[0]["counts"].sum() == nseq * niter and h[0]["counts"][0] == h[0]["counts"][-1] == 20
    np.testing.assert_allclose([h[1]["mean"], h[1]["bestfit"]], [90.0, 90.0], rtol=1e-6)
    assert h[1]["std"] < 1e-5
    prm = histogram_mean_params(pout, chi2, np.array([0.0, 7.0, 0.0], np.float32), invar)
    np.testing.assert_allclose(prm, [2.0, 7.0, np.pi / 2], rtol=1e-7)
