"""Visibility -> brightness post-processing (model_vis2bri_simple).

CPU: with the oracle's float64 model as visfun, the image sums to the
windowed visibility of the grid cell that the FFT treats as zero frequency
(only that term survives the sum over pixels; with the reference's
right-to-left U grid that cell is u = -du, v = 0), the uv and xy grids are
each other's Fourier duals, and the image of an unrotated source is
mirror-symmetric in y.
"""
import numpy as np

from cuda_mcmc_amd.postproc import model_vis2bri_simple
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
