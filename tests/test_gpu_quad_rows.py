"""The single-coordinate update of the quadratic-system step (k_iter_quad).

A proposal moves one coordinate x_t to x'_t = x_t + d, so the left side of
every equation moves by d (b_t + sum_j a_tj x_j + a_tt x'_t), where a_tj is
the coefficient of x_t x_j.  The kernel evaluates every equation in full once
per iteration, keeps sqrt(w_e) times each residual in registers and updates
them from a per-fit row table in LDS.

* the form follows the fit: ladders of at most 64 rungs, at most 8
  equations, non-negative weights; MHRE_QUAD_FULL=1 turns it off;
* from the same state its proposal chi^2 equals the full evaluation's to
  float rounding, and the decisions are the same;
* systems of 2, 4 and 8 unknowns, square and over-determined, run in lock
  step with the float64 oracle (reference update rule: gpu_mcmc.cu:361-626,
  model: examples/solvsys/model.cuh generalised to x^T Q x + b.x).
"""
import os

import numpy as np
import pytest

from cuda_mcmc_amd import quadsys, synthdata
from cuda_mcmc_amd.engine import Session
from oracle.chi2 import generic_chi2
from tests.test_gpu_mhre import _lockstep

pytestmark = pytest.mark.gpu


def _with_env(name, value, fn):
    old = os.environ.get(name)
    os.environ[name] = value
    try:
        return fn()
    finally:
        if old is None:
            del os.environ[name]
        else:
            os.environ[name] = old


def _fit(neq=8, n=8, nbeta=16, nseq=32, nburn=3, niter=0, seed=21, dseed=3, sigma=1.0, **kw):
    ds = synthdata.quadsys_dataset(nsys=1, neq=neq, n=n, seed=dseed)
    return quadsys.quad_fit(ds["Q"][0], ds["b"][0], ds["c"][0], nbeta=nbeta, nseq=nseq, nburn=nburn,
                            niter=niter, seed=seed, sigma=sigma, **kw)


def test_rows_form_follows_the_fit():
    neg = _fit(seed=4)
    neg.std2r[3] = -1.0
    fits = [_fit(), _fit(nbeta=70, nseq=2), _fit(neq=12, n=4), neg, _fit(neq=3, n=2), _fit(neq=6, n=4)]
    with Session(fits) as s:
        info = [s.fit_info(i) for i in range(len(fits))]
    assert [i["kernel"] for i in info] == ["quad"] * len(fits)
    assert [i["quad_rows"] for i in info] == [True, False, False, False, True, True]
    assert [i["exchange_in_step"] for i in info] == [True, False, True, True, True, True]

    def forced():
        with Session([_fit()]) as s:
            return s.fit_info(0)
    assert not _with_env("MHRE_QUAD_FULL", "1", forced)["quad_rows"]


def test_rows_form_matches_the_full_evaluation():
    # the same fit from the same initial state with both forms: the first
    # substep's proposals are the same, so their chi^2 differs only by float
    # rounding; later decisions stay the same outside ties
    def run():
        f = _fit(nbeta=16, nseq=64, nburn=3, seed=33)
        with Session([f], trace_iters=3) as s:
            rows = s.fit_info(0)["quad_rows"]
            s.run(3)
            s.download()
            tr = s.trace
        return rows, tr["dec"].copy(), tr["chi2"].copy()
    rows_a, dec_a, chi2_a = run()
    rows_b, dec_b, chi2_b = _with_env("MHRE_QUAD_FULL", "1", run)
    assert rows_a and not rows_b
    first = ((dec_a[0, 0] >> 8) & 1) == 1
    np.testing.assert_array_equal(dec_a[0, 0] & 0x1ff, dec_b[0, 0] & 0x1ff)
    assert first.sum() > 0.5 * first.size
    np.testing.assert_allclose(chi2_a[0, 0][first], chi2_b[0, 0][first], rtol=2e-5)
    assert np.mean(dec_a == dec_b) > 0.99


@pytest.mark.parametrize("neq,n,nbeta", [(2, 2, 16), (4, 4, 8), (6, 4, 16), (8, 8, 32)])
def test_lockstep_rows_form(neq, n, nbeta):
    fit = _fit(neq=neq, n=n, nbeta=nbeta, nseq=32, nburn=4, niter=3, seed=40 + neq + n, dseed=7 + n)
    with Session([_fit(neq=neq, n=n, nbeta=nbeta, nseq=32, nburn=4, niter=3)]) as s:
        assert s.fit_info(0)["quad_rows"]
    f = lambda P: generic_chi2("quadsys", P, fit.coor, fit.dat, fit.std2r)
    # fp32 residuals and Gaussian steps, as the full form (tests/test_gpu_generic.py)
    compared, counts = _lockstep(fit, f, 7, tie_rel=3e-5, edge_ulps=4, pcur_rtol=1e-6, pcur_step=1e-5)
    assert compared >= 7 * 32 * 0.5
    assert counts["ties"] <= 0.01 * counts["decisions"], counts


def test_lockstep_rows_form_fixed_and_angle_coordinates():
    # pdescr 0 holds coordinate 1 at its ptotal value (nprm 3 of nptot 4,
    # ivar/invar maps: the update row is the moved coordinate's, t = ivar[k]),
    # pdescr 2 makes coordinate 3 an angle (wrapped into [-pi, pi])
    from cuda_mcmc_amd.engine import Fit
    ds = synthdata.quadsys_dataset(nsys=1, neq=5, n=4, seed=17)
    Q = ds["Q"][0].astype(np.float32)
    b = ds["b"][0].astype(np.float32)
    coor = np.concatenate([Q.ravel(), b.ravel()]).astype(np.float32)
    dat = (-ds["c"][0]).astype(np.float32)
    pdescr = np.array([1, 0, 1, 2], np.int32)
    pmin = np.array([-5, -5, -5, -np.pi], np.float32)
    pmax = np.array([5, 5, 5, np.pi], np.float32)
    ptotal = np.array([0.0, 0.7, 0.0, 0.0], np.float32)

    def make():
        return Fit(pdescr, pmin, pmax, coor, dat, np.ones(5, np.float32), nbeta=16, nseq=32, nburn=4, niter=3,
                   model="quadsys", ptotal=ptotal, seed=91, record=True)
    fit = make()
    assert fit.nprm == 3 and fit.nptot == 4
    with Session([make()]) as s:
        assert s.fit_info(0)["quad_rows"]
    f = lambda P: generic_chi2("quadsys", P, fit.coor, fit.dat, fit.std2r)
    compared, counts = _lockstep(fit, f, 7, tie_rel=3e-5, edge_ulps=4, pcur_rtol=1e-6, pcur_step=1e-5)
    assert compared >= 7 * 32 * 0.5
    assert counts["ties"] <= 0.01 * counts["decisions"], counts
    assert np.all(fit.pcur[:, :, 1] == np.float32(0.7))
    assert np.all(np.abs(fit.pcur[:, :, 3]) <= np.float32(np.pi))
