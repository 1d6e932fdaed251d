"""The device approximation tables are what their generator makes, and their
fp32 errors (evaluated in fp32 on the host, as the kernels do) stay inside
the budget that keeps chi^2 within 1e-5 relative of the float64 oracle.

cuda_mcmc_amd/csrc/gen_approx.py fits Chebyshev/least-squares polynomials to
scipy's Bessel functions and numpy's sin/cos/arctan; approx_coeffs.h is its
output.  Independently of the generator's own checks, the table entries are
re-evaluated here against scipy/numpy in float64.
"""
import os
import re
import subprocess
import sys

import numpy as np
import pytest
from scipy.special import j1, jv

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "cuda_mcmc_amd", "csrc")
HDR = os.path.join(CSRC, "approx_coeffs.h")


def _arrays():
    txt = open(HDR).read()
    out = {}
    for name, body in re.findall(r"constexpr float (\w+)\[\d+\] = \{(.*?)\};", txt, re.S):
        out[name] = np.array([float(v.rstrip("f")) for v in re.findall(r"[-+0-9.e]+f", body)], np.float64)
    return out


def _horner32(c, t):
    acc = np.full_like(t, np.float32(c[-1]))
    for v in c[-2::-1]:
        acc = (acc * t + np.float32(v)).astype(np.float32)
    return acc


def test_header_is_generator_output():
    r = subprocess.run([sys.executable, os.path.join(CSRC, "gen_approx.py")], capture_output=True, text=True,
                       check=True)
    assert r.stdout == open(HDR).read()


def test_sin_cos_kernels_radians():
    a = _arrays()
    r = np.linspace(-np.pi / 2, np.pi / 2, 200001).astype(np.float32)
    r2 = (r * r).astype(np.float32)
    sn = (r + (r * r2).astype(np.float32) * _horner32(a["SIN_T"], r2)).astype(np.float32)
    cs = (np.float32(1) + r2 * _horner32(a["COS_T"], r2)).astype(np.float32)
    assert np.max(np.abs(sn - np.sin(r.astype(np.float64)))) < 1.5e-7
    assert np.max(np.abs(cs - np.cos(r.astype(np.float64)))) < 1.5e-7


def test_bessel_table_piecewise_quadratics():
    a = _arrays()
    tab = a["BESSEL_TAB"].reshape(-1, 8)
    assert tab.shape[0] == 257   # quarter-unit intervals of y over [0, 64]
    z = np.linspace(0.0, 256.0, 256001)   # z = 4y, the kernel's argument
    j = z.astype(np.int64)
    s = (z - j).astype(np.float32)
    B = _horner32_rows(tab[j, 0:3], s)
    C = _horner32_rows(tab[j, 3:6], s)
    y = z / 4.0
    x = np.sqrt(y)
    xs = np.maximum(x, 1e-3)
    Bt = np.where(x > 1e-3, j1(xs) / xs, 0.5 - y / 16.0)
    Ct = np.where(x > 1e-3, jv(2, xs) / (xs * xs), 0.125 - y / 96.0)
    assert np.max(np.abs(B - Bt)) < 6e-8
    assert np.max(np.abs(C - Ct)) < 2e-8
    # D(y) = J3(x)/x^3 for the gradient kernels, on the same intervals
    tabd = a["BESSEL_TABD"].reshape(-1, 4)
    assert tabd.shape[0] == tab.shape[0]
    Dv = _horner32_rows(tabd[j, 0:3], s)
    xd = np.maximum(x, 1e-2)
    Dt = np.where(x > 1e-2, jv(3, xd) / (xd * xd * xd), 1.0 / 48.0 - y / 768.0)
    assert np.max(np.abs(Dv - Dt)) < 5e-9
    # B(0) = 1/2 and C(0) = 1/8: the rho -> 0 limits of the reference model
    assert tab[0, 0] == pytest.approx(0.5, abs=1e-7) and tab[0, 3] == pytest.approx(0.125, abs=1e-7)


def _horner32_rows(c, s):
    acc = c[:, -1].astype(np.float32)
    for k in range(c.shape[1] - 2, -1, -1):
        acc = (acc * s + c[:, k].astype(np.float32)).astype(np.float32)
    return acc


def test_atan_kernel():
    a = _arrays()
    t = np.linspace(0.0, 1.0, 100001).astype(np.float32)
    p = (t * _horner32(a["ATAN_P"], (t * t).astype(np.float32))).astype(np.float32)
    assert np.max(np.abs(p - np.arctan(t.astype(np.float64)))) < 1.5e-7
