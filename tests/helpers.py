"""Shared builders for the tests (synthetic problems, oracle adapters)."""
import numpy as np

from cuda_mcmc_amd import synthdata
from cuda_mcmc_amd.engine import Fit


def eht_fit(d, nbeta, nseq, nburn=0, niter=0, seed=1, model="sgra9", model_data=False, **kw):
    idat = np.array([d["nvis"], nbeta * nseq * (d["nvis"] + d["ncph"]), 0], np.int32)
    if model == "sgra13":
        pdescr, pmin, pmax = sgra13_prior()
    else:
        pdescr, pmin, pmax = synthdata.EHT_PDESCR, synthdata.EHT_PMIN, synthdata.EHT_PMAX
    return Fit(pdescr, pmin, pmax, d["coor"], d["dat"], d["std2r"], nbeta=nbeta, nseq=nseq,
               nburn=nburn, niter=niter, model=model, imodel=1 if model == "sgra9" else 2,
               icoor=d["icoor"], idat=idat, npass=2, seed=seed, model_data=model_data, **kw)


def sgra13_prior():
    #            Zsp  Re   rq   ecc  fade gr   gax  aq   gq   alpha  beta   eta    th
    pmin = np.array([0.5, 20., 0.3, 0.0, 0.0, 0.0, 1.6, 0.5, 0.0, -np.pi, -np.pi, -np.pi, -np.pi], np.float32)
    pmax = np.array([3.0, 30., 0.9, 0.5, 1.0, 1.0, 7.5, 1.0, 0.5, np.pi, np.pi, np.pi, np.pi], np.float32)
    pdescr = np.array([1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2], np.int32)
    return pdescr, pmin, pmax


def random_params(rng, pmin, pmax, n):
    pmin = np.asarray(pmin, np.float64)
    pmax = np.asarray(pmax, np.float64)
    return (pmin + (pmax - pmin) * rng.uniform(size=(n, pmin.size))).astype(np.float32)


def oracle_eht_chi2(d, imodel=1):
    from oracle.chi2 import eht_chi2
    cpext = int(d.get("cpext", 0))

    def f(P):
        return eht_chi2(P, d["coor"], d["dat"], d["std2r"], d["icoor"], d["nvis"], d["ncph"], cpext, imodel)
    return f


def tie_tol(beta, c_new, c_cur, rel=2e-6):
    """Decision band inside which the device's fp32 model vs the oracle's fp64
    model may legitimately flip a decision.

    The device accumulates fp32 terms in fp64, so its chi^2 agrees with the
    oracle to ~1e-6 relative (the 1e-5 spec tolerance would flag almost every
    decision of a chain with chi^2 ~ 1e6-1e7 as a tie).  A decision that differs
    outside this band is a real discrepancy and fails the test.
    """
    return 0.5 * beta * rel * (np.abs(c_new) + np.abs(c_cur)) + 1e-6


def rosenbrock_moments():
    """First and second moments of p(x, y) ~ exp(-f^2 / 2) in the box
    [-5, 5]^2, f = 100 (y - x^2)^2 + (1 - x)^2, by quadrature on a grid
    (converged to about 1e-4)."""
    # the density lives along y = x^2 for x in about [-0.8, 2.8]; the prior
    # box cuts it at y = 5
    x = np.linspace(-1.5, 3.0, 1801)
    y = np.linspace(-1.0, 5.0, 4801)
    X, Y = np.meshgrid(x, y, indexing="ij")
    f = 100.0 * (Y - X * X) ** 2 + (1.0 - X) ** 2
    w = np.exp(-0.5 * f * f)
    z = w.sum()
    m = {"x": (w * X).sum() / z, "y": (w * Y).sum() / z}
    m.update(xx=(w * X * X).sum() / z, xy=(w * X * Y).sum() / z, yy=(w * Y * Y).sum() / z)
    return m
