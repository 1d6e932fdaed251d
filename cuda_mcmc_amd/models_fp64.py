"""Host-side float64 evaluation of the ring+Gaussian visibility model.

Used to synthesise observations and by host utilities (model images).  It is
written in the same reduced form as the device kernel: with x = 2 pi rho R and
y = x^2,

    J1(x)/rho                              = 2 pi R  * B(y),   B(y) = J1(x)/x
    pi R (J0-J2)(x)/rho^2 - J1(x)/rho^3    = -8 pi^3 R^3 C(y), C(y) = J2(x)/x^2

(J0 - J2 = 2 J0 - 2 J1/x by the recurrence J0 + J2 = 2 J1/x), which turns the
terms ree, ime, rei, imi of ``examples/sgra_model_fit/model.cuh:396-417`` into

    ree =  pi (h+h0) Re^2 B(ye)            ime = -2 pi^2 (h-h0) Re^3 U C(ye)
    rei =  pi (h+h0-d(h-h0)/Re) Ri^2 B(yi) imi = -2 pi^2 (h-h0) rq Ri^3 U C(yi)

with no 1/rho and no separate rho -> 0 branch (B(0) = 1/2, C(0) = 1/8 give the
reference's limits of model.cuh:411-416).
"""
import numpy as np
from scipy.special import j1, jv

UAS2NRAD = np.pi / 180.0 / 3600.0 * 1e3
GCOEF = 2.0 * (np.pi / (2.0 * np.sqrt(2.0 * np.log(2.0)))) ** 2


def bessel_B(y):
    """J1(x)/x with x = sqrt(y); series below x = 1e-3."""
    y = np.asarray(y, dtype=np.float64)
    x = np.sqrt(y)
    small = x < 1e-3
    xs = np.where(small, 1.0, x)
    return np.where(small, 0.5 - y / 16.0, j1(xs) / xs)


def bessel_C(y):
    """J2(x)/x^2 with x = sqrt(y); series below x = 1e-2."""
    y = np.asarray(y, dtype=np.float64)
    x = np.sqrt(y)
    small = x < 1e-2
    xs = np.where(small, 1.0, x)
    return np.where(small, 0.125 - y / 96.0 + y * y / 3072.0, jv(2, xs) / (xs * xs))


def sgra9_vis(u, v, p):
    """Complex visibility of the 9-parameter model; u, v in Glambda; p [9] or [..., 9]."""
    P = np.asarray(p, dtype=np.float64)
    cols = [P[..., i] if P.ndim == 1 else P[..., i:i + 1] for i in range(9)]
    Zsp, Re, rq, ecc, fade, gax, aq, gq, th = cols
    u = np.asarray(u, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    Rex = Re * UAS2NRAD
    Rin = rq * Rex
    d = ecc * (Rex - Rin)
    h = (2.0 / np.pi) / ((Rex ** 2 - Rin ** 2 * (1.0 + d / Rex)) * fade
                         + (Rex ** 2 - Rin ** 2 * (1.0 - d / Rex)))
    h0 = fade * h
    ct, st = np.cos(th), np.sin(th)
    U = u * ct + v * st
    V = -u * st + v * ct
    rho2 = U * U + V * V
    ye = (2.0 * np.pi * Rex) ** 2 * rho2
    yi = rq * rq * ye
    ree = np.pi * (h + h0) * Rex ** 2 * bessel_B(ye)
    ime = -2.0 * np.pi ** 2 * (h - h0) * Rex ** 3 * U * bessel_C(ye)
    rei = np.pi * (h + h0 - d * (h - h0) / Rex) * Rin ** 2 * bessel_B(yi)
    imi = -2.0 * np.pi ** 2 * (h - h0) * rq * Rin ** 3 * U * bessel_C(yi)
    Fr = (ree + 1j * ime) - np.exp(2j * np.pi * d * U) * (rei + 1j * imi)
    gsy = gax * Rex
    gsx = aq * gsy
    Gamp = np.exp(-GCOEF * ((U * gsx) ** 2 + (V * gsy) ** 2))
    Fg = Gamp * np.exp(-2j * np.pi * (Rin - d) * U)
    nog = (gax == 0.0) | (aq == 0.0) | (gq == 0.0)
    return Zsp * np.where(nog, Fr, (1.0 - gq) * Fr + gq * Fg)
