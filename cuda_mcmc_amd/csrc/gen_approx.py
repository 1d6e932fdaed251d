#!/usr/bin/env python3
"""Generate approx_coeffs.h: polynomial coefficients for the device math.

Every table is derived here from the mathematics (no third-party library
constants are copied):

BESSEL_B  B(y) = J1(x)/x, y = x^2 in [0, 64]: Chebyshev interpolation in
          t = y/32 - 1, converted to monomials in t (evaluated by Horner).
BESSEL_C  C(y) = J2(x)/x^2 on the same range and variable.
BESSEL_D  D(y) = J3(x)/x^3 on the same range and variable (derivative
          dC/dy = -D/2, used by the chi^2 gradient kernels).
BESSEL_TAB  piecewise cubics of B and C on the unit intervals [j, j+1) of y,
          j = 0..64, in s = y - j (least squares at Chebyshev nodes): 8
          floats per interval (B c0..c3, C c0..c3), read from LDS by the
          fused step kernel.
HANKEL    coefficients of Hankel's asymptotic expansion (Abramowitz & Stegun
          9.2.5, 9.2.9, 9.2.10) for J1 and J2, used for x > 8.
SIN/COS   least-squares odd / even polynomials for sin and cos on
          |r| <= pi/2 (the range after reduction by multiples of pi).
ATAN      odd polynomial t*P(t^2) for atan(t), t in [0, 1], Chebyshev fit.

The reference functions for the fits are SciPy's float64 Bessel functions.
Run:  python gen_approx.py > approx_coeffs.h
"""
import math
import numpy as np
from scipy.special import j1, jv


def B(y):
    x = np.sqrt(y)
    return np.where(x > 1e-6, j1(np.maximum(x, 1e-6)) / np.maximum(x, 1e-6), 0.5 - y / 16.0)


def C(y):
    x = np.sqrt(y)
    xs = np.maximum(x, 1e-3)
    return np.where(x > 1e-3, jv(2, xs) / (xs * xs), 0.125 - y / 96.0)


def D(y):
    x = np.sqrt(y)
    xs = np.maximum(x, 1e-2)
    return np.where(x > 1e-2, jv(3, xs) / (xs * xs * xs), 1.0 / 48.0 - y / 768.0)


def cheb_monomial(f, lo, hi, deg):
    n = 8 * deg
    k = np.arange(n)
    t = np.cos(np.pi * (k + 0.5) / n)
    y = lo + (t + 1) * (hi - lo) / 2
    c = np.polynomial.chebyshev.chebfit(t, f(y), deg)
    return np.polynomial.chebyshev.cheb2poly(c)


def horner32(m, t):
    t = np.float32(t)
    acc = np.float32(m[-1])
    for c in m[-2::-1]:
        acc = np.float32(np.float32(acc * t) + np.float32(c))
    return acc


def check32(m, f, lo, hi):
    y = np.linspace(lo, hi, 40001)
    t = (y / ((hi - lo) / 2) - 1).astype(np.float32)
    acc = np.full_like(t, np.float32(m[-1]))
    for c in m[-2::-1]:
        acc = (acc * t + np.float32(c)).astype(np.float32)
    return float(np.max(np.abs(acc.astype(np.float64) - f(y))))


def hankel_terms(n, kmax):
    """P and Q series coefficients a_k of A&S 9.2.9/9.2.10 in powers of 1/(8x)."""
    mu = 4.0 * n * n
    P, Q = [], []
    for k in range(kmax + 1):
        # term (mu-1)(mu-9)...(mu-(2k-1)^2) / (k! (8x)^k)
        num = 1.0
        for j in range(1, k + 1):
            num *= (mu - (2 * j - 1) ** 2)
        a = num / math.factorial(k)
        if k % 2 == 0:
            P.append(((-1) ** (k // 2)) * a)
        else:
            Q.append(((-1) ** ((k - 1) // 2)) * a)
    return P, Q


def odd_even_fit(f_odd, f_even, lim, n_odd, n_even):
    """sin-like r + r^3 P(r^2) and cos-like 1 + r^2 Q(r^2) least-squares fits on [-lim, lim]."""
    r = lim * np.cos(np.pi * (np.arange(2000) + 0.5) / 2000)
    r = r[r > 1e-3]
    s2 = r * r
    A = np.stack([r ** 3 * s2 ** k for k in range(n_odd)], 1)
    P = np.linalg.lstsq(A, f_odd(r) - r, rcond=None)[0]
    A = np.stack([s2 * s2 ** k for k in range(n_even)], 1)
    Q = np.linalg.lstsq(A, f_even(r) - 1.0, rcond=None)[0]
    return P, Q


def check_sincos32(P, Q, lim):
    r = np.linspace(-lim, lim, 200001).astype(np.float32)
    r2 = (r * r).astype(np.float32)
    acc = np.full_like(r2, np.float32(P[-1]))
    for c in P[-2::-1]:
        acc = (acc * r2 + np.float32(c)).astype(np.float32)
    sn = (r + (r * r2).astype(np.float32) * acc).astype(np.float32)
    acc = np.full_like(r2, np.float32(Q[-1]))
    for c in Q[-2::-1]:
        acc = (acc * r2 + np.float32(c)).astype(np.float32)
    cs = (np.float32(1.0) + r2 * acc).astype(np.float32)
    rd = r.astype(np.float64)
    return float(np.max(np.abs(sn - np.sin(rd)))), float(np.max(np.abs(cs - np.cos(rd))))


def atan_fit(deg_odd_terms):
    # atan(t) = t * P(t^2), t in [0, 1]; fit P in s = t^2 on [0,1]
    n = 400
    k = np.arange(n)
    t = np.cos(np.pi * (k + 0.5) / n)
    s = (t + 1) / 2
    tt = np.sqrt(s)
    g = np.where(tt > 0, np.arctan(tt) / np.where(tt > 0, tt, 1), 1.0)
    c = np.polynomial.chebyshev.chebfit(t, g, deg_odd_terms - 1)
    m = np.polynomial.chebyshev.cheb2poly(c)
    # m is in variable t' = 2s - 1; convert to monomials in s
    p = np.polynomial.polynomial.Polynomial(m)
    q = p(np.polynomial.polynomial.Polynomial([-1.0, 2.0]))
    return q.coef


def piecewise_cubic(f, nint, h=1.0, deg=3):
    """Per-interval least-squares cubics: [nint][deg+1] in s = (y - j h)/h."""
    s = np.cos(np.pi * (np.arange(64) + 0.5) / 64) * 0.5 + 0.5
    out = []
    for j in range(nint):
        out.append(np.polynomial.polynomial.polyfit(s, f((j + s) * h), deg))
    return np.array(out)


def check_piecewise32(c, f, h=1.0, upto=64.0):
    worst = 0.0
    ss = np.linspace(0.0, 1.0, 2001).astype(np.float32)
    for j in range(int(upto / h)):
        acc = np.full_like(ss, np.float32(c[j][-1]))
        for k in range(c.shape[1] - 2, -1, -1):
            acc = (acc * ss + np.float32(c[j][k])).astype(np.float32)
        worst = max(worst, float(np.max(np.abs(acc.astype(np.float64) - f((j + ss.astype(np.float64)) * h)))))
    return worst


def emit(name, coefs, comment):
    print("// %s" % comment)
    print("static constexpr int %s_N = %d;" % (name, len(coefs)))
    print("__device__ static constexpr float %s[%d] = {" % (name, len(coefs)))
    for c in coefs:
        print("    %.9ef," % c)
    print("};")
    print()


def main():
    print("// Generated by gen_approx.py -- do not edit by hand.")
    print("#pragma once")
    print()
    mB = cheb_monomial(B, 0.0, 64.0, 9)
    mC = cheb_monomial(C, 0.0, 64.0, 8)
    eB = check32(mB, B, 0.0, 64.0)
    eC = check32(mC, C, 0.0, 64.0)
    emit("BESSEL_B", mB, "J1(x)/x = sum c_k t^k, t = x^2/32 - 1, x in [0,8]; fp32 max abs err %.2e" % eB)
    emit("BESSEL_C", mC, "J2(x)/x^2 = sum c_k t^k, t = x^2/32 - 1, x in [0,8]; fp32 max abs err %.2e" % eC)
    nq = 257   # quarter-unit intervals of y covering [0, 64] (the last one only at y = 64)
    tb = piecewise_cubic(B, nq, h=0.25, deg=2)
    tc = piecewise_cubic(C, nq, h=0.25, deg=2)
    eb = check_piecewise32(tb, B, h=0.25)
    ec = check_piecewise32(tc, C, h=0.25)
    tab = np.concatenate([tb, tc, np.zeros((nq, 2))], axis=1).ravel()
    print("// B and C on y in [j/4, (j+1)/4), j = 0..256: c0 + s(c1 + s c2), s = 4y - j;")
    print("// 8 floats per interval (B0 B1 B2 C0 C1 C2 0 0); fp32 max abs err B %.2e, C %.2e on [0, 64]"
          % (eb, ec))
    print("static constexpr int BESSEL_TAB_N = %d;" % nq)
    print("__device__ static constexpr float BESSEL_TAB[%d] = {" % tab.size)
    for j in range(nq):
        print("    " + " ".join("%.9ef," % v for v in tab[8 * j:8 * j + 8]))
    print("};")
    print()
    td = piecewise_cubic(D, nq, h=0.25, deg=2)
    ed = check_piecewise32(td, D, h=0.25)
    tabd = np.concatenate([td, np.zeros((nq, 1))], axis=1).ravel()
    print("// D on the same intervals: D0 D1 D2 0 per interval (gradient kernels); fp32 max abs err %.2e" % ed)
    print("__device__ static constexpr float BESSEL_TABD[%d] = {" % tabd.size)
    for j in range(nq):
        print("    " + " ".join("%.9ef," % v for v in tabd[4 * j:4 * j + 4]))
    print("};")
    print()
    for n in (1, 2):
        P, Q = hankel_terms(n, 11)
        emit("HANKEL_P%d" % n, P, "P_%d(x) = sum a_k (8x)^(-2k)  (A&S 9.2.9)" % n)
        emit("HANKEL_Q%d" % n, Q, "Q_%d(x) = sum a_k (8x)^(-(2k+1))  (A&S 9.2.10)" % n)
    P, Q = odd_even_fit(np.sin, np.cos, np.pi / 2, 4, 5)
    es, ec = check_sincos32(P, Q, np.pi / 2 + 1e-6)
    emit("SIN_T", P, "sin r = r + r^3 sum c_k r^(2k), |r| <= pi/2 (least squares); fp32 max abs err %.2e" % es)
    emit("COS_T", Q, "cos r = 1 + r^2 sum c_k r^(2k), |r| <= pi/2 (least squares); fp32 max abs err %.2e" % ec)
    ma = atan_fit(10)
    t = np.linspace(0, 1, 20001).astype(np.float32)
    s = (t * t).astype(np.float32)
    acc = np.full_like(s, np.float32(ma[-1]))
    for c in ma[-2::-1]:
        acc = (acc * s + np.float32(c)).astype(np.float32)
    ea = float(np.max(np.abs((acc * t).astype(np.float64) - np.arctan(t.astype(np.float64)))))
    emit("ATAN_P", ma, "atan(t) = t * sum c_k (t^2)^k, t in [0,1]; fp32 max abs err %.2e" % ea)


if __name__ == "__main__":
    main()
