"""float64 torch restatement of the 9-parameter model and its chi^2, with
autograd, as an independent gradient reference for the device kernels.

Follows the literal formulas of examples/sgra_model_fit/model.cuh:322-462 (the
rho > 1e-7 branch with J0, J1 and Jn(2) at 2 pi rho Re and 2 pi rho Ri) and the
chi^2 terms of src/gpu_mcmc.cu:305-323.  J0 and J1 values come from SciPy
(float64, CPU; torch.special's float64 Bessel functions are only accurate to
~5e-7), with their derivatives J0' = -J1, J1' = J0 - J1/x supplied here and
Jn(2) = 2 J1/x - J0.  Test infrastructure only.
"""
import math

import numpy as np
import torch
from scipy import special

UAS2NRAD = math.pi / 180.0 / 3600.0 * 1e3
GCOEF = 2.0 * (math.pi / (2.0 * math.sqrt(2.0 * math.log(2.0)))) ** 2


def _sp(fn, x):
    return torch.as_tensor(fn(x.detach().cpu().numpy().astype(np.float64)), dtype=torch.float64)


class _J0(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        ctx.save_for_backward(x)
        return _sp(special.j0, x)

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        return -g * _sp(special.j1, x)


class _J1(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        ctx.save_for_backward(x)
        return _sp(special.j1, x)

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        return g * (_sp(special.j0, x) - _sp(special.j1, x) / x)


def j0(x):
    return _J0.apply(x)


def j1(x):
    return _J1.apply(x)


def j2(x):
    return 2.0 * j1(x) / x - j0(x)


def sgra9_vis(u, v, P):
    """(Re V, Im V) [n, nvis] for parameter rows P [n, 9] (float64 tensors)."""
    Zsp, Re, rq, ecc, fade, gax, aq, gq, th = [P[:, i:i + 1] for i in range(9)]
    Rex = Re * UAS2NRAD
    Rin = rq * Rex
    d = ecc * (Rex - Rin)
    h = (2.0 / math.pi) / ((Rex ** 2 - Rin ** 2 * (1.0 + d / Rex)) * fade + (Rex ** 2 - Rin ** 2 * (1.0 - d / Rex)))
    h0 = fade * h
    U = u * torch.cos(th) + v * torch.sin(th)
    V = -u * torch.sin(th) + v * torch.cos(th)
    rho = torch.sqrt(U ** 2 + V ** 2)
    pi4 = 4.0 * math.pi
    ae = 2.0 * math.pi * rho * Rex
    ree = 0.5 * (h + h0) * Rex * j1(ae) / rho
    ime = (h - h0) / pi4 * (math.pi * Rex * (j0(ae) - j2(ae)) / rho ** 2 - j1(ae) / rho ** 3) * U
    ai = 2.0 * math.pi * rho * Rin
    rei = 0.5 * (h + h0 - d * (h - h0) / Rex) * Rin * j1(ai) / rho
    imi = (h - h0) / pi4 * rq * (math.pi * Rin * (j0(ai) - j2(ai)) / rho ** 2 - j1(ai) / rho ** 3) * U
    ph = 2.0 * math.pi * d * U
    cd, sd = torch.cos(ph), torch.sin(ph)
    frx = ree - (rei * cd - imi * sd)
    fry = ime - (rei * sd + imi * cd)
    gsy = gax * Rex
    gsx = aq * gsy
    gamp = torch.exp(-GCOEF * ((U * gsx) ** 2 + (V * gsy) ** 2))
    gph = -2.0 * math.pi * (Rin - d) * U
    return Zsp * ((1.0 - gq) * frx + gq * gamp * torch.cos(gph)), Zsp * ((1.0 - gq) * fry + gq * gamp * torch.sin(gph))


def _wrap(x):
    return torch.atan2(torch.sin(x), torch.cos(x))


def eht_chi2(P, d):
    """chi^2 [n] of parameter rows P [n, 9] against a synthdata-style dataset d."""
    nvis, ncph = d["nvis"], d["ncph"]
    coor = torch.as_tensor(d["coor"], dtype=torch.float64)
    dat = torch.as_tensor(d["dat"], dtype=torch.float64)
    w = torch.as_tensor(d["std2r"], dtype=torch.float64)
    u, v = coor[0:2 * nvis:2], coor[1:2 * nvis:2]
    vx, vy = sgra9_vis(u, v, P)
    amp = torch.sqrt(vx ** 2 + vy ** 2)
    pha = torch.atan2(vy, vx)
    chi = ((dat[:nvis] - amp) ** 2 * w[:nvis]).sum(-1)
    if ncph:
        ic = torch.as_tensor(d["icoor"], dtype=torch.long)[:3 * ncph].reshape(ncph, 3)
        cph = pha[:, ic[:, 0]] + pha[:, ic[:, 1]] - pha[:, ic[:, 2]]
        chi = chi + (_wrap(dat[nvis:nvis + ncph] - cph) ** 2 * w[nvis:nvis + ncph]).sum(-1)
    return chi
