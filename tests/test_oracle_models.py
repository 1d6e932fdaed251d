"""The fp64 model restatements pinned by closed forms, invariants and brute force."""
import numpy as np
import pytest
from scipy.special import j1

from oracle.models import sgra_model9, sgra_model13, ring_image, UAS2NRAD, GCOEF, rosenbrock, solvsys, quadsys
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


# ------------------------------------------------------------ quadratic systems

def _solvsys_as_quadsys():
    """The reference solvsys system x^2 + y^2 = 1, -0.25 x + y = 0.5
    (examples/solvsys/model.cuh:51-54 with P1 = P2 = 2, K = 0.25) in the form
    x^T Q_e x + b_e . x: Q_0 = I, b_0 = 0; Q_1 = 0, b_1 = (-0.25, 1)."""
    Q = np.array([[[1.0, 0.0], [0.0, 1.0]], [[0.0, 0.0], [0.0, 0.0]]])
    b = np.array([[0.0, 0.0], [-0.25, 1.0]])
    return Q, b


def test_quadsys_reduces_to_the_reference_solvsys():
    Q, b = _solvsys_as_quadsys()
    P = rng.uniform(-3, 3, size=(200, 2))
    np.testing.assert_allclose(quadsys(P, Q, b), solvsys(P, [2, 2, 0.25]), rtol=1e-14, atol=1e-14)
    # closed-form roots (examples/solvsys/solvsys.py:59-62): the left-hand sides equal the
    # right-hand sides (1, 0.5), i.e. the residuals vanish
    for rx in ((-4.0 - np.sqrt(832.0)) / 34.0, (-4.0 + np.sqrt(832.0)) / 34.0):
        ry = 0.5 + 0.25 * rx
        np.testing.assert_allclose(quadsys([rx, ry], Q, b), [1.0, 0.5], atol=1e-14)
    np.testing.assert_allclose([(-4.0 - np.sqrt(832.0)) / 34.0, (-4.0 + np.sqrt(832.0)) / 34.0],
                               [-0.966, 0.7307], atol=5e-4)


def test_quadsys_hand_computed_nonsymmetric():
    # x^T Q x = 1*49 + 2*7*(-1) + 3*(-1)*7 + 4*1 = 18 ; b.x = 5*7 + 6*(-1) = 29
    Q = np.array([[[1.0, 2.0], [3.0, 4.0]]])
    b = np.array([[5.0, 6.0]])
    assert quadsys([7.0, -1.0], Q, b)[0] == 47.0
    # second equation with the transposed Q: the same quadratic form
    Q2 = np.concatenate([Q, Q.transpose(0, 2, 1)])
    b2 = np.concatenate([b, np.zeros((1, 2))])
    np.testing.assert_array_equal(quadsys([7.0, -1.0], Q2, b2), [47.0, 18.0])


def test_quadsys_antisymmetric_part_vanishes_and_b_is_linear():
    n, neq = 8, 8
    A = rng.normal(size=(neq, n, n))
    A = A - A.transpose(0, 2, 1)
    P = rng.uniform(-2, 2, size=(50, n))
    np.testing.assert_allclose(quadsys(P, A, np.zeros((neq, n))), 0.0, atol=1e-12)
    S = rng.normal(size=(neq, n, n))
    b = rng.normal(size=(neq, n))
    np.testing.assert_allclose(quadsys(P, S + A, b), quadsys(P, S, b), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(quadsys(P, np.zeros((neq, n, n)), b), P @ b.T, rtol=1e-13, atol=1e-13)
    # brute-force double sum for one point
    x = P[0]
    brute = [sum(S[e, i, j] * x[i] * x[j] for i in range(n) for j in range(n)) + sum(b[e] * x)
             for e in range(neq)]
    np.testing.assert_allclose(quadsys(x, S, b), brute, rtol=1e-12)


def test_generic_datm_quadsys_unpacks_the_fit_layout():
    """coor = [Q (neq*n*n), b (neq*n)] as packed by quadsys.quad_fit."""
    from oracle.chi2 import generic_datm, generic_chi2
    from cuda_mcmc_amd.quadsys import quad_fit
    Q, b = _solvsys_as_quadsys()
    fit = quad_fit(Q, b, np.array([-1.0, -0.5]), nbeta=2, nseq=2, nburn=1)
    P = rng.uniform(-3, 3, size=(20, 2))
    np.testing.assert_allclose(generic_datm("quadsys", P, fit.coor, 2), solvsys(P, [2, 2, 0.25]), rtol=1e-14)
    roots = np.array([[rx, 0.5 + 0.25 * rx] for rx in ((-4.0 - np.sqrt(832.0)) / 34.0, (-4.0 + np.sqrt(832.0)) / 34.0)])
    # fp32 coefficients: the residual at the exact roots is at float32 rounding level
    assert np.all(generic_chi2("quadsys", roots, fit.coor, fit.dat, fit.std2r) < 1e-6)


# ------------------------------------------- the reference's second fp64 mirror

def _xringaus_obsdatafit(ulam, vlam, vparam):
    """Point-by-point restatement of xringaus, src/obsdatafit.py:1476-1608 (rho -> 0
    limits below 1e-12, separate Ri == 0 branch), with math-module scalars."""
    import math
    from scipy.special import j0 as J0, j1 as J1, jn as JN
    Zsp, Re, rq, ecc, fade, gax, aq, gq, th = [float(t) for t in vparam]
    Re = math.radians(Re / 3600.0) * 1e3
    Ri = rq * Re
    d = ecc * (Re - Ri)
    gsy = gax * Re
    gsx = aq * gsy
    f = fade
    h = (2.0 / math.pi) / ((Re ** 2 - Ri ** 2 * (1.0 + d / Re)) * f + (Re ** 2 - Ri ** 2 * (1.0 - d / Re)))
    h0 = f * h
    out = []
    for ul, vl in zip(ulam, vlam):
        U = ul * math.cos(th) + vl * math.sin(th)
        V = -ul * math.sin(th) + vl * math.cos(th)
        rho = math.sqrt(U * U + V * V)
        if rho < 1e-12:
            ree = 0.5 * (h + h0) * math.pi * Re ** 2
            ime = -0.25 * math.pi ** 2 * (h - h0) * Re ** 3 * U
            rei = 0.5 * (h + h0 - d * (h - h0) / Re) * math.pi * Ri ** 2
            imi = -0.25 * math.pi ** 2 * (h - h0) * (Ri ** 4 / Re) * U
        else:
            a = 2.0 * math.pi * rho
            ree = 0.5 * (h + h0) * Re * J1(a * Re) / rho
            ime = (h - h0) / (4 * math.pi) * (math.pi * Re * (J0(a * Re) - JN(2, a * Re)) / rho ** 2
                                             - J1(a * Re) / rho ** 3) * U
            rei = 0.5 * (h + h0 - d * (h - h0) / Re) * Ri * J1(a * Ri) / rho
            imi = (h - h0) / (4 * math.pi) * (Ri / Re) * (math.pi * Ri * (J0(a * Ri) - JN(2, a * Ri)) / rho ** 2
                                                        - J1(a * Ri) / rho ** 3) * U
        Fex = complex(ree, ime)
        Fr = Fex - complex(math.cos(2 * math.pi * d * U), math.sin(2 * math.pi * d * U)) * complex(rei, imi) \
            if Ri != 0 else Fex
        if gax == 0.0 or aq == 0.0 or gq == 0.0:
            out.append(Zsp * Fr)
            continue
        gcoef = 2.0 * (math.pi / (2.0 * math.sqrt(2.0 * math.log(2.0)))) ** 2
        Gamp = math.exp(-gcoef * ((U * gsx) ** 2 + (V * gsy) ** 2))
        Gph = -2.0 * math.pi * (Ri - d) * U
        out.append(Zsp * ((1.0 - gq) * Fr + gq * Gamp * complex(math.cos(Gph), math.sin(Gph))))
    return np.array(out)


def test_oracle_matches_the_obsdatafit_mirror():
    """sgra_model9 (restating model.cuh:322-462) against a scalar restatement of the
    reference's independent fp64 mirror src/obsdatafit.py:1476-1608."""
    u, v = _uv(300)
    u[:2] = [0.0, 0.0]
    v[:2] = [0.0, 3e-13]
    for p in list(_rand_params(6)) + [np.array([1.2, 24.0, 0.0, 0.3, 0.5, 3.0, 0.8, 0.3, 0.2]),
                                      np.array([1.2, 24.0, 0.5, 0.3, 0.5, 3.0, 0.8, 0.0, 0.2])]:
        np.testing.assert_allclose(sgra_model9(u, v, p), _xringaus_obsdatafit(u, v, p), rtol=1e-10, atol=1e-13)
