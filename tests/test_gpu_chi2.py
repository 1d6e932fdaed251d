"""Native fused chi^2 (ring+Gaussian) against the float64 oracle.

Tolerances: chi^2 within 1e-5 relative; model amplitudes within 2e-6 of the
zero-spacing flux; closure phases and phases within 2e-5 rad away from the
+-pi branch cut.
"""
import numpy as np
import pytest

from cuda_mcmc_amd import mcmc_interf as mi
from cuda_mcmc_amd import synthdata
from oracle.chi2 import eht_chi2_terms
from tests.helpers import random_params, sgra13_prior

pytestmark = pytest.mark.gpu


def _calc(d, P, imodel, cpext=0):
    nb, ns = P.shape[0], P.shape[1]
    nvis, ncph = d["nvis"], d["ncph"]
    ndat = nvis + ncph
    C = nb * ns
    coor = d["coor"] if cpext == 0 else d["coor_ext"]
    idat = np.array([nvis, C * ndat, cpext], np.int32)
    datm = np.zeros(C * (ndat + nvis), np.float32)
    chi2m = np.zeros((nb, ns, ndat), np.float32)
    chi2c = np.zeros((nb, ns), np.float32)
    flag = np.ones((nb, ns), np.int32)
    icoor = d["icoor"] if cpext == 0 else np.zeros(1, np.int32)
    mi.calcmodchi2(coor, d["dat"], d["std2r"], icoor, idat, datm, chi2m, P, chi2c, flag, imodel,
                   coor.size, ndat, icoor.size, 3, datm.size, P.shape[2], nb, ns)
    return chi2c, datm, chi2m


@pytest.mark.parametrize("seed", [2, 11])
def test_sgra9_chi2_matches_oracle(seed):
    d = synthdata.ring_gauss_dataset(nvis=2000, ncph=600, seed=seed)
    rng = np.random.default_rng(seed)
    P = random_params(rng, synthdata.EHT_PMIN, synthdata.EHT_PMAX, 48).reshape(4, 12, 9)
    P[0, 0] = d["truth"].astype(np.float32)          # one chain at the truth (chi^2 ~ ndat)
    chi2c, datm, chi2m = _calc(d, P, 1)
    terms, head, pha = eht_chi2_terms(P.reshape(-1, 9), d["coor"], d["dat"], d["std2r"], d["icoor"],
                                      d["nvis"], d["ncph"], 0, 1)
    ref = terms.sum(-1).reshape(4, 12)
    np.testing.assert_allclose(chi2c, ref, rtol=1e-5)
    nvis, ndat, C = d["nvis"], d["nvis"] + d["ncph"], 48
    amp_gpu = datm[:C * ndat].reshape(C, ndat)[:, :nvis]
    zsp = P.reshape(-1, 9)[:, :1]
    np.testing.assert_allclose(amp_gpu, head[:, :nvis], atol=2e-6 * zsp.max(), rtol=2e-6)
    # phases: the fp32 absolute error of V (bounded like the amplitudes above,
    # 2e-6 Zsp) divided by |V|
    amp = head[:, :nvis]
    cph_gpu = datm[:C * ndat].reshape(C, ndat)[:, nvis:]
    dd = np.abs(np.angle(np.exp(1j * (cph_gpu - head[:, nvis:]))))
    ic = d["cpix"]
    inv = zsp / amp[:, ic[:, 0]] + zsp / amp[:, ic[:, 1]] + zsp / amp[:, ic[:, 2]]
    assert np.all(dd <= 2e-6 + 2e-6 * inv)
    pha_gpu = datm[C * ndat:].reshape(C, nvis)
    dp = np.abs(np.angle(np.exp(1j * (pha_gpu - pha))))
    assert np.all(dp <= 2e-6 + 2e-6 * zsp / amp)
    np.testing.assert_allclose(chi2m.reshape(C, ndat), terms, rtol=2e-4, atol=1e-6 * terms.max())


def test_sgra9_chi2_inner_radius_beyond_outer():
    """rq in (1, 1.5]: the inner ring's Bessel argument rq^2 ye is the larger one.

    The fused kernels take the table-only Bessel path for a tile only when both
    arguments stay within the table (y <= 64); with rq > 1 that bound comes from
    the inner ring.  Baselines up to 8 Glambda and Re up to 30 uas put ye just
    below 64, so rq^2 ye crosses it on the long-baseline tiles."""
    d = synthdata.ring_gauss_dataset(nvis=2000, ncph=600, seed=13)
    rng = np.random.default_rng(13)
    P = random_params(rng, synthdata.EHT_PMIN, synthdata.EHT_PMAX, 48).reshape(4, 12, 9)
    P[..., 1] = rng.uniform(27.0, 30.0, size=(4, 12))          # Re: ye close to 64 at 8 Glambda
    P[..., 2] = rng.uniform(1.25, 1.5, size=(4, 12))           # rq > 1
    P[0, 0, 2] = 1.05
    uv2 = d["coor"][0::2] ** 2 + d["coor"][1::2] ** 2
    x2 = (2 * np.pi * P[..., 1:2].astype(np.float64) * np.pi / 180 / 3600 * 1e3) ** 2 * uv2.max()
    assert np.all(x2 < 64) and np.sum(x2 * P[..., 2:3].astype(np.float64) ** 2 > 64) >= 40
    chi2c, _, _ = _calc(d, P, 1)
    terms, _, _ = eht_chi2_terms(P.reshape(-1, 9), d["coor"], d["dat"], d["std2r"], d["icoor"],
                                 d["nvis"], d["ncph"], 0, 1)
    np.testing.assert_allclose(chi2c, terms.sum(-1).reshape(4, 12), rtol=1e-5)


def test_sgra9_chi2_at_truth_is_data_sized():
    d = synthdata.ring_gauss_dataset(nvis=2000, ncph=600, seed=3)
    P = np.tile(d["truth"].astype(np.float32), (1, 1, 1))
    chi2c, _, _ = _calc(d, P, 1)
    terms, _, _ = eht_chi2_terms(P.reshape(-1, 9), d["coor"], d["dat"], d["std2r"], d["icoor"],
                                 d["nvis"], d["ncph"], 0, 1)
    np.testing.assert_allclose(chi2c.ravel(), terms.sum(-1), rtol=1e-5)
    assert 0.5 * 2600 < chi2c[0, 0] < 2.0 * 2600


def test_sgra13_chi2_matches_oracle():
    d = synthdata.ring_gauss_dataset(nvis=800, ncph=250, seed=5)
    pdescr, pmin, pmax = sgra13_prior()
    rng = np.random.default_rng(5)
    P = random_params(rng, pmin, pmax, 32).reshape(4, 8, 13)
    chi2c, datm, _ = _calc(d, P, 2)
    terms, head, _ = eht_chi2_terms(P.reshape(-1, 13), d["coor"], d["dat"], d["std2r"], d["icoor"],
                                    d["nvis"], d["ncph"], 0, 2)
    np.testing.assert_allclose(chi2c.ravel(), terms.sum(-1), rtol=1e-5)


def test_cpext1_matches_oracle():
    d = synthdata.ring_gauss_dataset(nvis=600, ncph=200, seed=6)
    # external closure coordinates: (u1,v1,u2,v2) of the three baselines' first two
    c = d["cpix"]
    uc = np.stack([d["u"][c[:, 0]], d["v"][c[:, 0]], d["u"][c[:, 1]], d["v"][c[:, 1]]], 1)
    d["coor_ext"] = np.concatenate([d["coor"], uc.ravel()]).astype(np.float32)
    rng = np.random.default_rng(6)
    P = random_params(rng, synthdata.EHT_PMIN, synthdata.EHT_PMAX, 16).reshape(2, 8, 9)
    chi2c, _, _ = _calc(d, P, 1, cpext=1)
    terms, _, _ = eht_chi2_terms(P.reshape(-1, 9), d["coor_ext"], d["dat"], d["std2r"], None,
                                 d["nvis"], d["ncph"], 1, 1)
    np.testing.assert_allclose(chi2c.ravel(), terms.sum(-1), rtol=1e-5)


def test_flag0_chain_sums_given_terms():
    """calc_chi2 sums every chi2m entry; flag-0 chains are not re-evaluated (gpu_mcmc.cu:291, :343)."""
    d = synthdata.ring_gauss_dataset(nvis=300, ncph=90, seed=8)
    rng = np.random.default_rng(8)
    P = random_params(rng, synthdata.EHT_PMIN, synthdata.EHT_PMAX, 4).reshape(1, 4, 9)
    nvis, ndat = d["nvis"], d["nvis"] + d["ncph"]
    idat = np.array([nvis, 4 * ndat, 0], np.int32)
    datm = np.zeros(4 * (ndat + nvis), np.float32)
    chi2m = np.full((1, 4, ndat), 0.5, np.float32)
    chi2c = np.zeros((1, 4), np.float32)
    flag = np.array([[1, 0, 1, 0]], np.int32)
    mi.calcmodchi2(d["coor"], d["dat"], d["std2r"], d["icoor"], idat, datm, chi2m, P, chi2c, flag, 1,
                   d["coor"].size, ndat, d["icoor"].size, 3, datm.size, 9, 1, 4)
    np.testing.assert_allclose(chi2c[0, [1, 3]], 0.5 * ndat, rtol=1e-6)
    assert np.all(chi2c[0, [0, 2]] != 0.5 * ndat)
