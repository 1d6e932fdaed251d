"""The two forms of the ring+Gaussian shift phases in the step kernel.

The hardware sine and cosine take their argument in revolutions and reduce it
themselves inside their input domain (|h| <= 256; tests/test_gpu_sincos.py
measures both forms), so the explicit reduction h - rint(h) is skipped for a
fit whose phases the host bounds by 64 revolutions over the prior box, the
chains' values and the data's longest baseline.  A fit that can exceed that
keeps the reduction; so does every fit under MHRE_REDUCE_PHASES=1.

* the choice follows the bound (a 13-parameter fit whose Gaussian sits
  hundreds of ring radii off centre has phases of ~700 revolutions);
* both forms give the same proposal chi^2 to 1e-6 relative;
* the large-phase fit runs in lock step with the float64 oracle.
"""
import os

import numpy as np
import pytest

from cuda_mcmc_amd import synthdata
from cuda_mcmc_amd.engine import Fit, Session
from tests.helpers import eht_fit, oracle_eht_chi2, sgra13_prior
from tests.test_gpu_mhre import _lockstep

pytestmark = pytest.mark.gpu


def _far_gaussian_fit(d, nb=8, ns=16, nburn=2, niter=2, seed=61, beta=None):
    # Gaussian offset gr in 400..600 ring radii (Re 20-30 uas, baselines to
    # ~8 G lambda): phases up to ~700 revolutions; a faint Gaussian (gq ~ 1e-3)
    # keeps the fp32 rounding of such phases out of the chi^2 comparison
    pdescr, pmin, pmax = sgra13_prior()
    pmin, pmax = pmin.copy(), pmax.copy()
    pmin[5], pmax[5] = 400.0, 600.0
    pmin[8], pmax[8] = 1e-3, 2e-3
    idat = np.array([d["nvis"], nb * ns * (d["nvis"] + d["ncph"]), 0], np.int32)
    return Fit(pdescr, pmin, pmax, d["coor"], d["dat"], d["std2r"], nbeta=nb, nseq=ns, nburn=nburn, niter=niter,
               model="sgra13", imodel=2, icoor=d["icoor"], idat=idat, npass=2, seed=seed, beta=beta)


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


def test_phase_form_follows_the_bound():
    d = synthdata.ring_gauss_dataset(nvis=400, ncph=120, seed=62)
    with Session([eht_fit(d, 4, 8, seed=1), eht_fit(d, 4, 8, seed=2, model="sgra13"), _far_gaussian_fit(d)]) as s:
        info = [s.fit_info(i) for i in range(3)]
    assert [i["kernel"] for i in info] == ["sgra9", "sgra13", "sgra13"]
    assert [i["hw_phase"] for i in info] == [True, True, False]
    assert all(i["exchange_in_step"] for i in info)

    def forced():
        with Session([eht_fit(d, 4, 8, seed=1)]) as s:
            return s.fit_info(0)
    assert not _with_env("MHRE_REDUCE_PHASES", "1", forced)["hw_phase"]


@pytest.mark.parametrize("model", ["sgra9", "sgra13"])
def test_unreduced_phases_give_the_same_chi2(model):
    # one iteration from the same initial state and the same draws: every
    # proposal is the same, so its chi^2 may differ only by the hardware
    # sine's rounding of the two arguments
    d = synthdata.ring_gauss_dataset(nvis=1500, ncph=450, seed=63)

    def run():
        f = eht_fit(d, 8, 16, nburn=1, niter=0, seed=64, model=model)
        with Session([f], trace_iters=1) as s:
            hw = s.fit_info(0)["hw_phase"]
            s.run(1)
            s.download()
            tr = s.trace
        return hw, tr["dec"][0].copy(), tr["chi2"][0].copy()
    hw, dec_a, chi2_a = run()
    red, dec_b, chi2_b = _with_env("MHRE_REDUCE_PHASES", "1", run)
    assert hw and not red
    # the first substep's proposals are the same in both runs
    np.testing.assert_array_equal(dec_a[0] & 0x1ff, dec_b[0] & 0x1ff)
    inprior = ((dec_a[0] >> 8) & 1) == 1
    np.testing.assert_allclose(chi2_a[0][inprior], chi2_b[0][inprior], rtol=1e-6)


def test_lockstep_with_phases_beyond_the_hardware_domain():
    # random starts put chi^2 at 1e5-2e7, where the 2e-6-relative tie band is
    # tens of nats at beta = 1; a hot ladder (beta 1e-3 .. 1e-4) keeps the
    # band to a fraction of a percent of the decisions (0.3 % in the oracle)
    d = synthdata.ring_gauss_dataset(nvis=300, ncph=90, seed=65)
    beta = np.geomspace(1e-3, 1e-4, 4).astype(np.float32)
    fit = _far_gaussian_fit(d, nb=4, beta=beta)
    with Session([_far_gaussian_fit(d, nb=4, beta=beta)]) as s:
        assert not s.fit_info(0)["hw_phase"]
    compared, counts = _lockstep(fit, oracle_eht_chi2(d, imodel=2), 4)
    assert compared >= 4 * 16 * 0.5
    assert counts["ties"] <= 0.02 * counts["decisions"], counts
