"""The fp64 model restatements pinned by closed forms, invariants and brute force."""
import numpy as np
import pytest
from scipy.special import j1

from oracle.models import sgra_model9, sgra_model13, ring_image, UAS2NRAD, GCOEF, rosenbrock, solvsys
from cuda_mcmc_amd.models_fp64 import sgra9_vis
from cuda_mcmc_amd.synthdata import EHT_PMIN, EHT_PMAX

rng = np.random.default_rng(0)


def _rand_params(n):
    return EHT_PMIN + (EHT_PMAX - EHT_PMIN) * rng.uniform(size=(n, 9))


def _uv(n=500, r=9.0):
    return rng.uniform(-r, r, n), rng.uniform(-r, r, n)


def test_zero_spacing_is_total_flux():
    for p in _rand_params(20):
        V = sgra_model9(np.array([0.0]), np.array([0.0]), p)
        assert abs(V[0] - p[0]) < 1e-12 * p[0]


def test_uniform_disk_is_airy():
    # fade = 1 (flat brightness), rq = 0 (no hole), no Gaussian: V = Zsp 2 J1(x)/x, x = 2 pi rho R
    p = np.array([1.7, 25.0, 0.0, 0.0, 1.0, 3.0, 0.7, 0.0, 0.4])
    u, v = _uv()
    x = 2 * np.pi * np.hypot(u, v) * p[1] * UAS2NRAD
    np.testing.assert_allclose(sgra_model9(u, v, p), p[0] * 2 * j1(x) / x, rtol=1e-12, atol=1e-14)


def test_pure_gaussian():
    p = np.array([2.1, 22.0, 0.5, 0.2, 0.3, 4.0, 0.6, 1.0, 1.1])
    u, v = _uv()
    th = p[8]
    U = u * np.cos(th) + v * np.sin(th)
    V = -u * np.sin(th) + v * np.cos(th)
    Rex = p[1] * UAS2NRAD
    gsy = p[5] * Rex
    gsx = p[6] * gsy
    amp = p[0] * np.exp(-GCOEF * ((U * gsx) ** 2 + (V * gsy) ** 2))
    np.testing.assert_allclose(np.abs(sgra_model9(u, v, p)), amp, rtol=1e-12, atol=1e-300)


def test_hermitian_linearity_and_half_turn():
    u, v = _uv()
    for p in _rand_params(5):
        V = sgra_model9(u, v, p)
        np.testing.assert_allclose(sgra_model9(-u, -v, p), np.conj(V), rtol=1e-11, atol=1e-13)
        q = p.copy()
        q[0] *= 2.5
        np.testing.assert_allclose(sgra_model9(u, v, q), 2.5 * V, rtol=1e-12, atol=1e-14)
        q = p.copy()
        q[8] += np.pi
        np.testing.assert_allclose(sgra_model9(u, v, q), np.conj(V), rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("p", [[1.5, 25.0, 0.6, 0.3, 0.4, 3.0, 0.7, 0.2, 0.5],
                               [0.8, 21.0, 0.8, 0.45, 0.05, 2.0, 0.9, 0.35, -2.3]])
def test_direct_fourier_sum_of_the_image(p):
    """The analytic visibility is the Fourier transform of the slashed-ring image."""
    p = np.array(p)
    L, n = 0.6, 1400
    x = np.linspace(-L, L, n)
    dx = x[1] - x[0]
    X, Y = np.meshgrid(x, x, indexing="ij")
    img = ring_image(X, Y, p)
    assert abs(img.sum() * dx * dx - p[0]) < 2e-3 * p[0]
    for (u, v) in [(1.0, 2.0), (3.0, -1.0), (-2.0, 0.5), (0.3, -4.0)]:
        Vb = (img * np.exp(-2j * np.pi * (u * X + v * Y))).sum() * dx * dx
        assert abs(Vb - sgra_model9(u, v, p)) < 3e-3 * p[0]


def test_reduced_form_matches_literal_restatement():
    """The host model in B(y), C(y) form (product) against the literal formulas (oracle)."""
    u, v = _uv(2000)
    u[:3] = [0.0, 1e-9, 0.0]
    v[:3] = [0.0, 0.0, 2e-8]
    P = _rand_params(30)
    np.testing.assert_allclose(sgra9_vis(u, v, P), sgra_model9(u, v, P), rtol=1e-9, atol=1e-12)


def test_model13_reduces_to_model9():
    u, v = _uv()
    for p in _rand_params(5):
        Zsp, Re, rq, ecc, fade, gax, aq, gq, th = p
        gr = rq - ecc * (1 - rq)          # Gaussian centre Rin - d in units of Re
        q = np.array([Zsp, Re, rq, ecc, fade, gr, gax, aq, gq, th, th, th, th])
        np.testing.assert_allclose(sgra_model13(u, v, q), sgra_model9(u, v, p), rtol=1e-10, atol=1e-12)
        assert abs(sgra_model13(np.zeros(1), np.zeros(1), q)[0] - Zsp) < 1e-12


def test_stacked_parameters_broadcast():
    u, v = _uv(50)
    P = _rand_params(4)
    stacked = sgra_model9(u, v, P)
    for i in range(4):
        np.testing.assert_allclose(stacked[i], sgra_model9(u, v, P[i]), rtol=1e-13)


def test_toy_models_known_answers():
    assert rosenbrock([1.0, 1.0]) == 0.0
    assert rosenbrock([0.0, 0.0]) == 1.0
    # solvsys roots of x^2 + y^2 = 1, -0.25 x + y = 0.5 (examples/solvsys/solvsys.py:59-62)
    for rx in ((-4.0 - np.sqrt(832.0)) / 34.0, (-4.0 + np.sqrt(832.0)) / 34.0):
        ry = np.sqrt(1 - rx ** 2)
        np.testing.assert_allclose(solvsys([rx, ry], [2, 2, 0.25]), [1.0, 0.5], atol=1e-12)
