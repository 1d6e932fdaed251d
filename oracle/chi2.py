"""Closures and chi^2 terms of the reference (oracle side, float64).

eht_datm        model data of ``examples/sgra_model_fit/model.cuh:9-115``: pass 0
                amplitudes (head) and phases (tail); pass 1 closure phases from
                three stored phases (cpExt=0) or three fresh model evaluations
                at (u1,v1), (u2,v2), (-(u1+u2), -(v1+v2)) (cpExt=1).
eht_chi2_terms  ``src/gpu_mcmc.cu:305-323``: dif = dat - datm, wrapped with
                atan2(sin, cos) for closure terms (model returns 2), squared
                and weighted by std2r.
generic_chi2    the same for the toy models (Rosenbrock, solvsys, quadsys).

Parity: pinned by hand-computed three-station cases and by the identity
wrap(a + b - c) = arg(e^{ia} e^{ib} e^{-ic}) (tests/test_oracle_chi2.py).
"""
import numpy as np
from .models import sgra_model9, sgra_model13, rosenbrock, solvsys, quadsys


def wrap(x):
    """atan2(sin x, cos x) -- the reference's angle correction (gpu_mcmc.cu:316)."""
    return np.arctan2(np.sin(x), np.cos(x))


def _model(imodel):
    if imodel == 1:
        return sgra_model9
    if imodel == 2:
        return sgra_model13
    raise ValueError("imodel must be 1 (9-param) or 2 (13-param), got %r" % (imodel,))


def eht_datm(P, coor, icoor, nvis, ncph, cpext=0, imodel=1):
    """Model amplitudes [..., nvis], closure phases [..., ncph], phases [..., nvis]."""
    f = _model(imodel)
    coor = np.asarray(coor, dtype=np.float64)
    u = coor[0:2 * nvis:2]
    v = coor[1:2 * nvis:2]
    V = f(u, v, P)
    amp = np.abs(V)
    pha = np.angle(V)
    if ncph == 0:
        cph = np.zeros(amp.shape[:-1] + (0,))
    elif cpext == 0:
        ic = np.asarray(icoor, dtype=np.int64)[:3 * ncph].reshape(ncph, 3)
        cph = wrap(pha[..., ic[:, 0]] + pha[..., ic[:, 1]] - pha[..., ic[:, 2]])
    else:
        uc = coor[2 * nvis:2 * nvis + 4 * ncph].reshape(ncph, 4)
        u1, v1, u2, v2 = uc[:, 0], uc[:, 1], uc[:, 2], uc[:, 3]
        p1 = np.angle(f(u1, v1, P))
        p2 = np.angle(f(u2, v2, P))
        p3 = np.angle(f(-(u1 + u2), -(v1 + v2), P))
        cph = wrap(p1 + p2 - p3)
    return amp, cph, pha


def eht_chi2_terms(P, coor, dat, std2r, icoor, nvis, ncph, cpext=0, imodel=1):
    """Per-datum chi^2 terms [..., nvis+ncph] and the model data head/tail."""
    amp, cph, pha = eht_datm(P, coor, icoor, nvis, ncph, cpext, imodel)
    dat = np.asarray(dat, dtype=np.float64)
    w = np.asarray(std2r, dtype=np.float64)
    ta = (dat[:nvis] - amp) ** 2 * w[:nvis]
    tc = wrap(dat[nvis:nvis + ncph] - cph) ** 2 * w[nvis:nvis + ncph]
    return np.concatenate([ta, tc], axis=-1), np.concatenate([amp, cph], axis=-1), pha


def eht_chi2(P, coor, dat, std2r, icoor, nvis, ncph, cpext=0, imodel=1):
    return eht_chi2_terms(P, coor, dat, std2r, icoor, nvis, ncph, cpext, imodel)[0].sum(-1)


def generic_datm(model, P, coor, ndat):
    """Model data for the toy plugins.  model in {'rosenbrock','solvsys','quadsys'}.

    For 'quadsys' coor holds Q (neq*n*n) then b (neq*n) with neq = ndat and
    n = P.shape[-1].
    """
    P = np.asarray(P, dtype=np.float64)
    if model == "rosenbrock":
        return rosenbrock(P)[..., None]
    if model == "solvsys":
        return solvsys(P, coor)
    if model == "quadsys":
        n = P.shape[-1]
        coor = np.asarray(coor, dtype=np.float64)
        Q = coor[:ndat * n * n].reshape(ndat, n, n)
        b = coor[ndat * n * n:ndat * n * n + ndat * n].reshape(ndat, n)
        return quadsys(P, Q, b)
    if model == "template":
        return np.zeros(P.shape[:-1] + (ndat,))
    raise ValueError(model)


def generic_chi2(model, P, coor, dat, std2r):
    dat = np.asarray(dat, dtype=np.float64)
    datm = generic_datm(model, P, coor, dat.size)
    return ((dat - datm) ** 2 * np.asarray(std2r, dtype=np.float64)).sum(-1)
