"""The Python 3 driver classes on the GPU, against the oracle.

* Mcgpu_Sgra on a committed uv fixture (examples/sgra_model_fit data, 323
  visibilities, closures built from the phases): a short fit runs; attribute
  shapes follow src/imgpu.py; calcmodchi2 of any parameter rows equals the
  oracle's chi^2 (1e-5 relative), and of the best sample equals the recorded
  chi^2 of that sample;
* calcmodel on a uv grid equals the oracle's 9- and 13-parameter visibility
  amplitudes and phases;
* the brightness image of model_vis2bri_simple from GPU visibilities equals
  the one from the oracle's (2e-5 of the peak).
"""
import os

import numpy as np
import pytest

from cuda_mcmc_amd import imgpu
from cuda_mcmc_amd.uvdata import snr_closure_sigma
from oracle import models
from oracle.chi2 import eht_chi2
from tests.helpers import random_params, sgra13_prior

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
UV = os.path.join(GOLD, "000008_1_1_000snd_uvdata.txt")

# prior of examples/sgra_model_fit/model_obsfit.py (9 parameters)
PDESCR = np.array((1, 1, 1, 1, 1, 1, 1, 1, 2), np.int32)
PTOTAL = np.array((2.4, 0., 0., 0., 1., 0., 0., 0., 0.), np.float32)
BND = np.array(((0.01, 5.0), (20., 60.), (0., 0.99), (0., 1.), (0., 1.), (0.1, 2.), (0.1, 1.5), (0., 1.),
                (-np.pi, np.pi)), np.float32)


def _fitter(**kw):
    pherr = np.loadtxt(os.path.join(GOLD, "pherr_vs_snr.txt"))
    probe = imgpu.Mcgpu_Sgra(uvfile=UV, pdescr=PDESCR, ptotal=PTOTAL, pmint=BND[:, 0], pmaxt=BND[:, 1],
                             ecph=1.0, **kw)
    pe = pherr[:, 1] if pherr.ndim == 2 else pherr
    ecph = snr_closure_sigma(probe.amp, probe.eamp, probe.cpix, pe)
    return imgpu.Mcgpu_Sgra(uvfile=UV, pdescr=PDESCR, ptotal=PTOTAL, pmint=BND[:, 0], pmaxt=BND[:, 1],
                            ecph=ecph, **kw)


def test_mcgpu_sgra_fit_and_calcmodchi2():
    mc = _fitter(nbeta=8, nseq=16, nburn=300, niter=200, seed=7)
    assert mc.error == ""
    assert mc.nvis == 323 and mc.ncph > 0
    mc.burnin_and_search()
    nprm, nseq, niter, nbeta = int(mc.nprm), int(mc.nseq), int(mc.niter), int(mc.nbeta)
    assert mc.pout.shape == (nprm, nseq * niter)
    assert mc.chi2.shape == (nseq * niter,)
    assert mc.pout_4d.shape == (nprm, nbeta, nseq, niter)
    assert mc.rate_acpt.shape == (nprm, nbeta)
    assert mc.mvicp.shape == (nbeta, nseq, int(mc.ndat)) and mc.mpha.shape == (nbeta, nseq, int(mc.nvis))
    assert np.all(np.isfinite(mc.chi2))
    # chi^2 of the best sample, recomputed
    i = int(np.argmin(mc.chi2))
    best = PTOTAL.copy()
    best[mc.ivar] = mc.pout[:, i]
    c2, mvicp, mpha, chi2m, datm = mc.calcmodchi2(best)
    np.testing.assert_allclose(c2, mc.chi2[i], rtol=1e-5)
    assert mvicp.shape == (int(mc.ndat),) and mpha.shape == (int(mc.nvis),)
    # random rows against the oracle
    P = random_params(np.random.default_rng(3), BND[:, 0], BND[:, 1], 12)
    c2, _, _, _, _ = mc.calcmodchi2(P)
    ref = eht_chi2(P.astype(np.float64), mc.coor, mc.dat, mc.std2r, mc.icoor, int(mc.nvis), int(mc.ncph))
    np.testing.assert_allclose(c2.ravel(), ref, rtol=1e-5)


@pytest.mark.parametrize("imodel", [1, 2])
def test_calcmodel_matches_oracle(imodel):
    mc = _fitter(nbeta=2, nseq=2, nburn=1, niter=1)
    g = np.linspace(-6.0, 6.0, 9, dtype=np.float32)
    ul, vl = np.meshgrid(g, g + 0.37)
    if imodel == 1:
        P = random_params(np.random.default_rng(5), BND[:, 0], BND[:, 1], 4)
        V = models.sgra_model9(ul.ravel(), vl.ravel(), P.astype(np.float64))
    else:
        _, lo, hi = sgra13_prior()
        P = random_params(np.random.default_rng(6), lo, hi, 4)
        V = models.sgra_model13(ul.ravel(), vl.ravel(), P.astype(np.float64))
    mamp, mpha = mc.calcmodel(ul, vl, P, imodel=imodel)
    assert mamp.shape == (4,) + ul.shape
    amp = np.abs(V).reshape(mamp.shape)
    np.testing.assert_allclose(mamp, amp, rtol=2e-5, atol=2e-6 * np.abs(P[:, :1, None]).max())
    dph = np.angle(np.exp(1j * (mpha - np.angle(V).reshape(mpha.shape))))
    assert np.all(np.abs(dph) < 2e-5 + 2e-6 * np.abs(P[:, 0, None, None]) / np.maximum(amp, 1e-12))


def test_vis2bri_gpu_matches_oracle_image():
    from cuda_mcmc_amd.postproc import gpu_visfun, model_vis2bri_simple
    prm = np.array([2.0, 25.0, 0.6, 0.3, 0.4, 3.0, 0.8, 0.2, 0.7], np.float32)

    def oracle_vf(U, V, p):
        return models.sgra_model9(U.ravel(), V.ravel(), p.astype(np.float64)).reshape(U.shape)

    _, Bg, *_ = model_vis2bri_simple(gpu_visfun(1), prm, 20.0, 128)
    _, Bo, *_ = model_vis2bri_simple(oracle_vf, prm, 20.0, 128)
    np.testing.assert_allclose(Bg.real, Bo.real, rtol=0, atol=2e-5 * np.abs(Bo.real).max())
