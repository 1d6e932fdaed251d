"""Failure paths that the host can force, and the index-redraw loops.

* An unknown imodel: the reference's model() returns -1 there
  (examples/sgra_model_fit/model.cuh:63) and calc_chi2_terms sums it as data
  (src/gpu_mcmc.cu:301); every entry point here returns MHRE_EMODEL instead.
* A negative amplitude weight: the native kernels store amplitudes as
  (d sqrt(w), sqrt(w)), so std2r < 0 (no inverse variance) is MHRE_EINVAL at
  session creation, and the library works afterwards.
* The pick redraw: the coordinate pick and the exchange pair pick are redrawn
  when the uniform rounds to n (src/gpu_mcmc.cu:369-372, :549-552).  The keys
  of tests/golden/redraw_keys.json make the first Philox word 0xFFFFFFFF at
  chain 0 / sequence 0 of iteration 0, so the device must take the retry
  counter's word exactly as the oracle does.
"""
import json
import os

import numpy as np
import pytest

from cuda_mcmc_amd import _lib, synthdata
from cuda_mcmc_amd import mcmc_interf as mi
from cuda_mcmc_amd.engine import Fit, Session
from oracle import philox as px
from tests.helpers import eht_fit, oracle_eht_chi2
from tests.test_gpu_mhre import _lockstep

pytestmark = pytest.mark.gpu

MHRE_EMODEL = -3
FX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "redraw_keys.json")))


def _small():
    return synthdata.ring_gauss_dataset(nvis=200, ncph=60, nsta=6, seed=41)


def _eht_arrays(d, nb, ns, imodel):
    ndat = d["nvis"] + d["ncph"]
    C = nb * ns
    idat = np.array([d["nvis"], C * ndat, 0], np.int32)
    datm = np.zeros(C * (ndat + d["nvis"]), np.float32)
    chi2m = np.zeros((nb, ns, ndat), np.float32)
    chi2c = np.zeros((nb, ns), np.float32)
    flag = np.ones((nb, ns), np.int32)
    P = np.tile(d["truth"].astype(np.float32), (nb, ns, 1))
    return idat, datm, chi2m, chi2c, flag, P


@pytest.mark.parametrize("imodel", [0, 3, -1])
@pytest.mark.parametrize("generic", [False, True])
def test_unknown_imodel_is_an_error(imodel, generic, monkeypatch):
    if generic:
        monkeypatch.setenv("MHRE_FORCE_GENERIC", "1")
    d = _small()
    ndat = d["nvis"] + d["ncph"]
    idat, datm, chi2m, chi2c, flag, P = _eht_arrays(d, 2, 3, imodel)
    with pytest.raises(_lib.EngineError, match=r"\(%d\)" % MHRE_EMODEL):
        mi.calcmodchi2(d["coor"], d["dat"], d["std2r"], d["icoor"], idat, datm, chi2m, P, chi2c, flag, imodel,
                       d["coor"].size, ndat, d["icoor"].size, 3, datm.size, 9, 2, 3)
    with pytest.raises(_lib.EngineError, match=r"\(%d\)" % MHRE_EMODEL):
        mi.calcmodel(d["coor"], idat, datm, P, imodel, d["coor"].size, d["nvis"], 3, datm.size, 9, 2, 3)
    f = Fit(synthdata.EHT_PDESCR, synthdata.EHT_PMIN, synthdata.EHT_PMAX, d["coor"], d["dat"], d["std2r"],
            nbeta=2, nseq=3, nburn=1, niter=1, imodel=imodel, icoor=d["icoor"], idat=idat, npass=2)
    with pytest.raises(_lib.EngineError, match=r"\(%d\)" % MHRE_EMODEL):
        mi.mcmcuda(f.coor, f.dat, f.std2r, f.icoor, f.idat, np.zeros(2 * 3 * (ndat + d["nvis"]), np.float32),
                   np.zeros((2, 3, ndat), np.float32), f.pdescr, f.ivar, f.invar, f.ptotal, f.pmint, f.pmaxt,
                   f.pout, f.tout, f.chi2, f.chi2c, f.tcur, f.flag, f.n_acpt, f.n_exch, f.n_cnt, f.n_hist,
                   f.beta, f.pstp, f.ptent, f.ptentn, f.pcur, f.rndst, 1, imodel, f.coor.size, ndat,
                   f.icoor.size, 3, 2 * 3 * (ndat + d["nvis"]), 9, 9, f.nadj, 2, 2, 3, 1, 1)
    # the library still works afterwards
    idat, datm, chi2m, chi2c, flag, P = _eht_arrays(d, 2, 3, 1)
    mi.calcmodchi2(d["coor"], d["dat"], d["std2r"], d["icoor"], idat, datm, chi2m, P, chi2c, flag, 1,
                   d["coor"].size, ndat, d["icoor"].size, 3, datm.size, 9, 2, 3)
    assert np.all(np.isfinite(chi2c)) and np.all(chi2c > 0)


def test_negative_amplitude_weight_is_an_error():
    d = _small()
    std2r = d["std2r"].copy()
    std2r[3] = -1.0
    idat = np.array([d["nvis"], 2 * 3 * (d["nvis"] + d["ncph"]), 0], np.int32)
    f = Fit(synthdata.EHT_PDESCR, synthdata.EHT_PMIN, synthdata.EHT_PMAX, d["coor"], d["dat"], std2r,
            nbeta=2, nseq=3, nburn=1, niter=1, model="sgra9", imodel=1, icoor=d["icoor"], idat=idat, npass=2)
    with pytest.raises(_lib.EngineError, match=r"std2r\[3\].*non-negative"):
        Session([f])
    g = Fit(synthdata.EHT_PDESCR, synthdata.EHT_PMIN, synthdata.EHT_PMAX, d["coor"], d["dat"], d["std2r"],
            nbeta=2, nseq=3, nburn=1, niter=1, model="sgra9", imodel=1, icoor=d["icoor"], idat=idat, npass=2)
    with Session([g]) as s:
        s.run(2)
        s.download()
    assert np.all(np.isfinite(g.chi2c))


def _seed(e):
    return e["key0"] | (e["key1"] << 32)


def test_step_pick_redraw_matches_oracle():
    e = FX["step_pick"]
    assert e["counter"] == [0, 0, 0, 0]
    d = _small()
    fit = eht_fit(d, nbeta=4, nseq=16, nburn=1, niter=1, seed=_seed(e))
    w0 = px.philox4x32(0, 0, 0, 0, e["key0"], e["key1"])[0]
    assert px.pick(w0, fit.nprm) == fit.nprm          # chain 0, substep 0 must redraw
    compared, _ = _lockstep(fit, oracle_eht_chi2(d), 2)
    assert compared >= 2 * 16 * 0.5
    retry = px.pick(px.philox4x32(0, 0, 0, 1, e["key0"], e["key1"])[0], fit.nprm)
    assert retry < fit.nprm


def test_exchange_pick_redraw_matches_oracle():
    e = FX["exchange_pick"]
    assert e["counter"] == [0, 0, 0, px.EXCH]
    d = _small()
    fit = eht_fit(d, nbeta=32, nseq=4, nburn=1, niter=1, seed=_seed(e))
    w0 = px.philox4x32(0, 0, 0, px.EXCH, e["key0"], e["key1"])[0]
    assert px.pick(w0, fit.nbeta - 1) == fit.nbeta - 1   # sequence 0, attempt 0 must redraw
    with Session([fit], trace_iters=1) as s:
        s.run(1)
        s.download()
        ex = s.trace["ex"][0]
    want = px.pick(px.philox4x32(0, 0, 0, px.EXCH + 1, e["key0"], e["key1"])[0], fit.nbeta - 1)
    assert (ex[0, 0] & 0xFFFF) == want < fit.nbeta - 1
    # and the whole iteration in lock step with the oracle
    fit = eht_fit(d, nbeta=32, nseq=4, nburn=1, niter=1, seed=_seed(e))
    compared, _ = _lockstep(fit, oracle_eht_chi2(d), 2)
    assert compared >= 1
