"""fp64 restatements of the reference's model() plugins (oracle side).

Functions
---------
sgra_model9(u, v, prm)   restates ``examples/sgra_model_fit/model.cuh:322-462``
                         (9-parameter "xringaus"); the same formulas appear in
                         the reference's fp64 mirror ``src/obsdatafit.py:1476-1608``.
sgra_model13(u, v, prm)  restates ``examples/sgra_model_fit/model.cuh:120-315``
                         (13-parameter "xringaus2").
ring_image(x, y, prm)    image-plane brightness whose Fourier transform the
                         9-parameter formulas are (derived from them; used only
                         for the brute-force Fourier-sum check).
rosenbrock(p)            ``examples/search_Rosenbrock/model.cuh:40-41``.
solvsys(p, coor)         ``examples/solvsys/model.cuh:51-54`` (datm for id 0,1).
quadsys(p, Q, b)         x^T Q_e x + b_e . x per equation (the large solvsys form).

Parity
------
sgra_model9   pinned: zero-spacing flux V(0,0)=Zsp, the uniform-disk (Airy)
              limit, the pure-Gaussian limit, Hermitian symmetry, linearity in
              Zsp, the th -> th+pi conjugation, and a direct Fourier sum of
              ring_image on a fine pixel grid (tests/test_oracle_models.py).
sgra_model13  pinned: reduces to sgra_model9 when eta = th and the 13-param
              Gaussian is placed where the 9-param one is; V(0,0)=Zsp;
              Hermitian symmetry.
rosenbrock, solvsys  pinned by their known roots/minima.
quadsys       pinned by its reduction to the reference's solvsys system
              (Q_0 = I, b_1 = (-0.25, 1)) and that system's closed-form roots,
              a hand-computed non-symmetric case, the vanishing of the
              antisymmetric part of Q and a brute-force double sum.
sgra_model9 is also compared with a point-by-point restatement of the
reference's second fp64 mirror, src/obsdatafit.py:1476-1608 (tests).

Units follow the reference: u, v in Gigalambda; radii and Gaussian axes in
microarcseconds (``model.cuh:366``).
"""
import numpy as np
from scipy.special import j0, j1, jn

UAS2NRAD = np.pi / 180.0 / 3600.0 * 1e3
GCOEF = 2.0 * (np.pi / (2.0 * np.sqrt(2.0 * np.log(2.0)))) ** 2

SGRA9_NAMES = ("Zsp", "Re", "rq", "ecc", "fade", "gax", "aq", "gq", "th")
SGRA13_NAMES = ("Zsp", "Re", "rq", "ecc", "fade", "gr", "gax", "aq", "gq",
                "alpha", "beta", "eta", "th")


def _ring_terms(U, rho, Rex, Rin, rq, h, h0, d):
    """ree, ime, rei, imi of model.cuh:389-417 (rho>1e-7 branch and its limit).

    Arguments broadcast (parameters may be arrays over chains).  When Rin == 0
    the general expressions give rei = imi = 0, which is the reference's
    ``Rin != 0`` branch (model.cuh:399, :424).
    """
    pi = np.pi
    pi4 = 4.0 * np.pi
    big = rho > 1e-7
    r = np.where(big, rho, 1.0)
    aRex = 2.0 * pi * r * Rex
    ree = 0.5 * (h + h0) * Rex * j1(aRex) / r
    ime = (h - h0) / pi4 * (pi * Rex * (j0(aRex) - jn(2, aRex)) / r ** 2
                            - j1(aRex) / r ** 3) * U
    ree = np.where(big, ree, 0.5 * (h + h0) * pi * Rex ** 2)
    ime = np.where(big, ime, -0.25 * pi ** 2 * (h - h0) * Rex ** 3 * U)
    aRin = 2.0 * pi * r * Rin
    rei = 0.5 * (h + h0 - d * (h - h0) / Rex) * Rin * j1(aRin) / r
    imi = (h - h0) / pi4 * rq * (pi * Rin * (j0(aRin) - jn(2, aRin)) / r ** 2
                                 - j1(aRin) / r ** 3) * U
    rei = np.where(big, rei, 0.5 * (h + h0 - d * (h - h0) / Rex) * pi * Rin ** 2)
    imi = np.where(big, imi, -0.25 * pi ** 2 * (h - h0) * (Rin ** 4 / Rex) * U)
    return ree, ime, rei, imi


def _ring_vis(U, V, Rex, Rin, rq, fade, d):
    """Fr = Fex - Fin of model.cuh:373-438 (complex, unit flux)."""
    h = (2.0 / np.pi) / ((Rex ** 2 - Rin ** 2 * (1.0 + d / Rex)) * fade
                         + (Rex ** 2 - Rin ** 2 * (1.0 - d / Rex)))
    h0 = fade * h
    rho = np.sqrt(U ** 2 + V ** 2)
    ree, ime, rei, imi = _ring_terms(U, rho, Rex, Rin, rq, h, h0, d)
    dis = np.exp(1j * 2.0 * np.pi * d * U)
    return (ree + 1j * ime) - dis * (rei + 1j * imi)


def _split(prm, n):
    """Parameter columns as float64 arrays shaped to broadcast against data.

    prm is [n] (one vector) or [..., n] (a stack); each returned column has a
    trailing axis of length 1 in the stacked case so it broadcasts against
    u, v of shape [nvis].
    """
    P = np.asarray(prm, dtype=np.float64)
    if P.ndim == 1:
        return [P[i] for i in range(n)]
    return [P[..., i:i + 1] for i in range(n)]


def sgra_model9(u, v, prm):
    """Complex visibility of the 9-parameter ring+Gaussian (model.cuh:322-462).

    u, v: [nvis] in Gigalambda.  prm: [9] -> [nvis], or [..., 9] -> [..., nvis].
    """
    Zsp, Re, rq, ecc, fade, gax, aq, gq, th = _split(prm, 9)
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    Rex = Re * UAS2NRAD
    Rin = rq * Rex
    d = ecc * (Rex - Rin)
    gsy = gax * Rex
    gsx = aq * gsy
    U = u * np.cos(th) + v * np.sin(th)
    V = -u * np.sin(th) + v * np.cos(th)
    Fr = _ring_vis(U, V, Rex, Rin, rq, fade, d)
    gshift = Rin - d
    Gamp = np.exp(-GCOEF * ((U * gsx) ** 2 + (V * gsy) ** 2))
    Gph = -2.0 * np.pi * gshift * U
    Fg = Gamp * np.exp(1j * Gph)
    nogauss = (gax == 0.0) | (aq == 0.0) | (gq == 0.0)
    return Zsp * np.where(nogauss, Fr, (1.0 - gq) * Fr + gq * Fg)


def sgra_model13(u, v, prm):
    """Complex visibility of the 13-parameter model (model.cuh:120-315).

    When the Gaussian is unspecified and gq == 1 the reference reads an
    uninitialised Fg (model.cuh:188-210); here that case returns 0.
    """
    (Zsp, Re, rq, ecc, fade, gr, gax, aq, gq, alpha, beta, eta,
     th) = _split(prm, 13)
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    Rex = Re * UAS2NRAD
    Rin = rq * Rex
    gauss = (gax != 0.0) & (aq != 0.0) & (gq != 0.0)
    gsy = gax * Rex
    gsx = aq * gsy
    Ub = u * np.cos(beta) + v * np.sin(beta)
    Vb = -u * np.sin(beta) + v * np.cos(beta)
    gshift = gr * Rex * (np.cos(alpha) * u + np.sin(alpha) * v)
    Gamp = np.exp(-GCOEF * ((Ub * gsx) ** 2 + (Vb * gsy) ** 2))
    Fg = np.where(gauss, Gamp * np.exp(1j * (-2.0 * np.pi * gshift)), 0.0)
    D = ecc * (Rex - Rin)
    d = D * np.cos(eta - th)
    U = u * np.cos(th) + v * np.sin(th)
    V = -u * np.sin(th) + v * np.cos(th)
    Fr = _ring_vis(U, V, Rex, Rin, rq, fade, d)
    mix = np.where(gauss, (1.0 - gq) * Fr + gq * Fg, Fr)
    return Zsp * np.where(gq == 1.0, Fg, mix)


def ring_image(x, y, prm):
    """Brightness (per nrad^2, unit total flux before Zsp) of the 9-param model.

    Derived from the Fourier formulas of model.cuh:396-458: the outer disk of
    radius Re carries a brightness ramp g(x') = (h+h0)/2 + (h-h0) x'/(2 Re)
    (h0 at x'=-Re, h at x'=+Re), the inner disk of radius Ri centred at
    x'=-d is removed, and an elliptical Gaussian (FWHM gsx along x', gsy along
    y') sits at x'=Ri-d.  (x', y') is (x, y) rotated by th, matching the uv
    rotation U = u cos th + v sin th.  x, y in nanoradians.
    """
    Zsp, Re, rq, ecc, fade, gax, aq, gq, th = [float(t) for t in prm[:9]]
    Rex = Re * UAS2NRAD
    Rin = rq * Rex
    d = ecc * (Rex - Rin)
    h = (2.0 / np.pi) / ((Rex ** 2 - Rin ** 2 * (1.0 + d / Rex)) * fade
                         + (Rex ** 2 - Rin ** 2 * (1.0 - d / Rex)))
    h0 = fade * h
    xp = x * np.cos(th) + y * np.sin(th)
    yp = -x * np.sin(th) + y * np.cos(th)
    g = 0.5 * (h + h0) + (h - h0) * xp / (2.0 * Rex)
    ring = g * ((xp ** 2 + yp ** 2 <= Rex ** 2).astype(np.float64)
                - ((xp + d) ** 2 + yp ** 2 <= Rin ** 2).astype(np.float64))
    if gax == 0.0 or aq == 0.0 or gq == 0.0:
        return Zsp * ring
    gsy = gax * Rex
    gsx = aq * gsy
    sx = gsx / (2.0 * np.sqrt(2.0 * np.log(2.0)))
    sy = gsy / (2.0 * np.sqrt(2.0 * np.log(2.0)))
    gshift = Rin - d
    gauss = np.exp(-0.5 * ((xp - gshift) / sx) ** 2 - 0.5 * (yp / sy) ** 2) / (2 * np.pi * sx * sy)
    return Zsp * ((1.0 - gq) * ring + gq * gauss)


def rosenbrock(p):
    """(1-x)^2 + 100 (y-x^2)^2 (examples/search_Rosenbrock/model.cuh:40-41)."""
    p = np.asarray(p, dtype=np.float64)
    x, y = p[..., 0], p[..., 1]
    return (1.0 - x) ** 2 + 100.0 * (y - x ** 2) ** 2


def solvsys(p, coor):
    """datm for id 0 and 1 of examples/solvsys/model.cuh:51-54.

    id 0: x**P1 + y**P2 with P1=coor[0], P2=coor[1]; id 1: -K*x + y, K=coor[2].
    Returns an array [..., 2].
    """
    p = np.asarray(p, dtype=np.float64)
    x, y = p[..., 0], p[..., 1]
    P1, P2, K = float(coor[0]), float(coor[1]), float(coor[2])
    return np.stack([np.power(x, P1) + np.power(y, P2), -K * x + y], axis=-1)


def quadsys(p, Q, b):
    """LHS x^T Q_e x + b_e . x of a system of quadratic equations.

    p: [..., n]; Q: [neq, n, n]; b: [neq, n]  ->  [..., neq]
    """
    p = np.asarray(p, dtype=np.float64)
    return np.einsum("...i,eij,...j->...e", p, Q, p) + np.einsum("...i,ei->...e", p, b)
