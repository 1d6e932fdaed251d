"""MH-RE decision parity: device engine vs the float64 oracle, in lock step.

Every iteration the oracle restarts from the state the GPU downloaded, runs
the same iteration (same Philox counters) and the two are compared:

* random coordinate picks and exchange pair picks: identical everywhere;
* prior checks, accept and swap decisions: identical except where the
  oracle's decision margin |log u - x| is within the chi^2 tolerance band
  (tests.helpers.tie_tol: 2e-6 relative chi^2 -> 0.5 beta 2e-6 (chi2 + chi2c));
* chain state of every sequence without a near-tie: counters, histories,
  temperature labels and step sizes bit-identical, parameters identical,
  chi^2 within 1e-5 relative.
"""
import numpy as np
import pytest

from cuda_mcmc_amd import synthdata
from cuda_mcmc_amd.engine import Session
from oracle import mhre
from tests.helpers import eht_fit, oracle_eht_chi2, tie_tol

pytestmark = pytest.mark.gpu

STATE = ("pcur", "chi2c", "pstp", "n_cnt", "n_acpt", "n_hist", "n_exch", "tcur", "flag", "ptent", "ptentn")


def _problem(fit, chi2_fn):
    return mhre.Problem(chi2_fn, fit.pdescr, fit.pmint, fit.pmaxt, fit.beta, fit.nseq, ptotal=fit.ptotal,
                        nadj=fit.nadj, seed=fit.seed, adapt_target=fit.adapt_target)


def _state_from(pb, fit):
    st = mhre.State(pb, niter=fit.niter)
    for k in STATE + ("pout", "tout", "chi2"):
        setattr(st, k, getattr(fit, k).copy())
    return st


def _compare_iteration(pb, fit, st, tr_mh, tr_ex, gpu_dec, gpu_ex, it, tie_rel=2e-6, counts=None, edge_ulps=0,
                       edge_step=0.0):
    """Compares one iteration's decisions; counts (a dict) accumulates the
    decisions made ("decisions") and those inside the tie band ("ties").
    edge_ulps, edge_step > 0: the device's proposal may differ from the
    oracle's by edge_ulps ulp plus edge_step times the step size, so a prior
    check whose proposal lies that close to a prior edge is a tie too."""
    B, S, P = fit.nbeta, fit.nseq, fit.nprm
    C = B * S
    bad_chain = np.zeros(C, bool)
    ntie = 0
    ndec = 0
    for s in range(P):
        o = tr_mh[s]
        g = gpu_dec[s]
        gk = g & 0xFF
        gin = (g >> 8) & 1
        gacc = (g >> 9) & 1
        live = ~bad_chain
        assert np.array_equal(gk[live], o["k"][live]), "coordinate pick differs (iteration %d, substep %d)" % (it, s)
        if edge_ulps or edge_step:
            p = o["p"].astype(np.float32)
            band = edge_ulps * np.spacing(np.abs(p)).astype(np.float64) + edge_step * np.abs(o["sd"])
            edge = ((np.abs(p.astype(np.float64) - o["lo"]) <= band) |
                    (np.abs(p.astype(np.float64) - o["hi"]) <= band))
            flip = live & edge & (gin.astype(bool) != o["inprior"])
            ntie += int(np.sum(flip))
            bad_chain |= flip
            live = ~bad_chain
        assert np.array_equal(gin[live].astype(bool), o["inprior"][live]), "prior check differs"
        beta = fit.beta[np.arange(C) // S]
        tie = o["inprior"] & (np.abs(o["margin"]) < tie_tol(beta, o["chi2"], o["chi2"], tie_rel))
        ntie += int(np.sum(tie & live))
        ndec += int(np.sum(o["inprior"] & live))
        diff = live & (gacc.astype(bool) != o["accept"])
        assert not np.any(diff & ~tie), "accept decision differs away from a tie (iteration %d, substep %d)" % (it, s)
        bad_chain |= tie
    bad_seq = bad_chain.reshape(B, S).any(axis=0)
    o = tr_ex[0]
    if B > 1:
        for a in range(B):
            gk = gpu_ex[a] & 0xFFFF
            gacc = (gpu_ex[a] >> 16) & 1
            ok = ~bad_seq
            assert np.array_equal(gk[ok], o["k"][a][ok]), "exchange pair pick differs"
            kk = o["k"][a]
            db = np.abs(fit.beta[kk + 1].astype(np.float64) - fit.beta[kk])
            c0 = np.abs(st.chi2c[kk, np.arange(S)])
            tie = np.abs(o["margin"][a]) < tie_tol(db, c0, c0, tie_rel)
            d = ok & (gacc.astype(bool) != o["accept"][a])
            assert not np.any(d & ~tie), "swap decision differs away from a tie"
            ntie += int(np.sum(tie & ok))
            ndec += int(np.sum(ok))
            bad_seq |= tie
    if counts is not None:
        counts["decisions"] = counts.get("decisions", 0) + ndec
        counts["ties"] = counts.get("ties", 0) + ntie
    return bad_seq, ntie


def _check_state(fit, st, ok, it, tie_rel=2e-6, pcur_rtol=1e-6, pcur_step=0.0, chi2_rtol=None):
    """The device state of every sequence without a near-tie equals the oracle's.
    pcur_step > 0: parameters may also differ by that multiple of the chain's
    largest step size (a kernel that draws the Gaussian step in fp32)."""
    for name in ("n_cnt", "n_acpt", "n_hist", "tcur", "n_exch", "flag", "ptentn"):
        g = getattr(fit, name)
        o = getattr(st, name)
        assert np.array_equal(g[..., ok], o[..., ok]), "%s differs at iteration %d" % (name, it)
    np.testing.assert_array_equal(fit.pstp[..., ok], st.pstp[..., ok])
    g, o = fit.pcur[:, ok].astype(np.float64), st.pcur[:, ok].astype(np.float64)
    tol = pcur_rtol * np.abs(o) + pcur_step * fit.pstp[:, :, ok].max(axis=0)[..., None]
    bad = np.abs(g - o) > tol
    assert not bad.any(), "pcur differs at iteration %d: %d elements, worst |d| / tol %.3g" % (
        it, bad.sum(), (np.abs(g - o) / np.maximum(tol, 1e-30)).max())
    np.testing.assert_allclose(fit.chi2c[:, ok], st.chi2c[:, ok], rtol=chi2_rtol or max(1e-5, 5 * tie_rel))
    if it >= fit.nburn and fit.pout is not None:
        itr = it - fit.nburn
        g, o = fit.pout[:, :, ok, itr].astype(np.float64), st.pout[:, :, ok, itr].astype(np.float64)
        bad = np.abs(g - o) > 1e-6 * np.abs(o) + pcur_step * fit.pstp[:, :, ok]
        assert not bad.any(), "pout differs at iteration %d: %d elements" % (it, bad.sum())
        np.testing.assert_array_equal(fit.tout[:, ok, itr], st.tout[:, ok, itr])


def _lockstep_batch(fits, chi2_fns, niters, tie_rel=2e-6, counts=None, edge_ulps=0, pcur_rtol=1e-6,
                    pcur_step=0.0, chi2_rtol=None, **session_options):
    """Lock step of a session of one or more fits: every iteration each fit's
    oracle restarts from the state the device downloaded.  Returns the number
    of (fit, sequence) states compared over all iterations.  pcur_step > 0
    allows parameters to differ by that many step sizes and widens the
    prior-edge tie band by as much."""
    pbs = [_problem(f, c) for f, c in zip(fits, chi2_fns)]
    compared = 0
    with Session(fits, trace_iters=niters, **session_options) as ses:
        ses.download()
        traces = ses.trace if isinstance(ses.trace, list) else [ses.trace]
        for it in range(niters):
            states, otr = [], []
            for fit, pb in zip(fits, pbs):
                st = _state_from(pb, fit)
                trace = {}
                burn = it < fit.nburn
                mhre.iteration(pb, st, it, burn, itr_out=None if burn else it - fit.nburn, trace=trace)
                states.append(st)
                otr.append(trace)
            ses.run(1)
            ses.download()
            for fit, pb, st, trace, tv in zip(fits, pbs, states, otr, traces):
                bad_seq, _ = _compare_iteration(pb, fit, st, trace["mh"], trace["ex"], tv["dec"][it], tv["ex"][it],
                                                it, tie_rel, counts, edge_ulps, pcur_step)
                ok = ~bad_seq
                compared += int(ok.sum())
                _check_state(fit, st, ok, it, tie_rel, pcur_rtol, pcur_step, chi2_rtol)
    return compared


def _lockstep(fit, chi2_fn, niters, tie_rel=2e-6, edge_ulps=0, pcur_rtol=1e-6, pcur_step=0.0, chi2_rtol=None,
              **session_options):
    counts = {}
    compared = _lockstep_batch([fit], [chi2_fn], niters, tie_rel, counts, edge_ulps, pcur_rtol, pcur_step, chi2_rtol,
                               **session_options)
    return compared, counts


def test_init_bit_exact():
    d = synthdata.ring_gauss_dataset(nvis=400, ncph=120, seed=21)
    fit = eht_fit(d, nbeta=8, nseq=8, nburn=0, niter=0, seed=12345)
    with Session([fit]) as ses:
        ses.download()
    pb = _problem(fit, oracle_eht_chi2(d))
    st = mhre.State(pb, niter=0)
    mhre.init_pcur(pb, st)
    np.testing.assert_array_equal(fit.pcur, st.pcur)
    np.testing.assert_allclose(fit.chi2c, st.chi2c, rtol=1e-5)
    assert np.all(fit.flag == 1)


def test_lockstep_eht_burnin_and_sampling():
    d = synthdata.ring_gauss_dataset(nvis=500, ncph=150, seed=22)
    fit = eht_fit(d, nbeta=8, nseq=16, nburn=3, niter=3, seed=77)
    compared, counts = _lockstep(fit, oracle_eht_chi2(d), 6)
    assert compared >= 6 * 16 * 0.5
    assert counts["ties"] <= 0.02 * counts["decisions"], counts


def test_lockstep_adaptation_threshold_and_short_history():
    # nadj smaller than the run so the history ring wraps; adapt target 0.25
    d = synthdata.ring_gauss_dataset(nvis=300, ncph=90, seed=23)
    fit = eht_fit(d, nbeta=4, nseq=16, nburn=8, niter=1, seed=5, nadj=3, adapt_target=0.25)
    compared, _ = _lockstep(fit, oracle_eht_chi2(d), 9)
    assert compared >= 9 * 16 * 0.5


@pytest.mark.parametrize("nbeta,nseq", [(1, 5), (3, 7), (5, 13), (40, 3), (70, 2)])
def test_lockstep_exchange_ladder_shapes(nbeta, nseq):
    # ladders of up to 32 rungs share a wave between 64 / 2^ceil(log2 nbeta)
    # sequences (here with a partly filled last wave; one rung: 64 sequences
    # per wave and no swap attempts), 33..64 rungs take a wave per sequence
    # and longer ladders the serial exchange
    d = synthdata.ring_gauss_dataset(nvis=200, ncph=60, seed=24)
    fit = eht_fit(d, nbeta=nbeta, nseq=nseq, nburn=2, niter=2, seed=31 + nbeta)
    compared, counts = _lockstep(fit, oracle_eht_chi2(d), 4)
    # a long ladder makes (nprm + 1) B decisions per sequence and iteration, so
    # a sequence is more often left out of the state comparison for one
    # decision inside a tie band; the decisions themselves were all compared
    # and the tie band holds few of them (random starts: chi^2 ~ 1e5-1e7)
    assert counts["ties"] <= 0.01 * counts["decisions"], counts
    assert compared >= 1


def test_lockstep_fixed_parameters():
    # pdescr 0: ecc and the Gaussian axis ratio are held at their ptotal values
    # (nprm 7 of nptot 9, ivar/invar maps), the orientation stays an angle
    from cuda_mcmc_amd.engine import Fit
    d = synthdata.ring_gauss_dataset(nvis=300, ncph=90, seed=25)
    pdescr = synthdata.EHT_PDESCR.copy()
    pdescr[3] = 0
    pdescr[6] = 0
    ptotal = (0.5 * (synthdata.EHT_PMIN + synthdata.EHT_PMAX)).astype(np.float32)
    ptotal[3], ptotal[6] = 0.21, 0.8
    idat = np.array([d["nvis"], 8 * 16 * (d["nvis"] + d["ncph"]), 0], np.int32)
    fit = Fit(pdescr, synthdata.EHT_PMIN, synthdata.EHT_PMAX, d["coor"], d["dat"], d["std2r"], nbeta=8, nseq=16,
              nburn=3, niter=3, model="sgra9", imodel=1, icoor=d["icoor"], idat=idat, ptotal=ptotal, npass=2,
              seed=66)
    assert fit.nprm == 7 and fit.nptot == 9
    compared, _ = _lockstep(fit, oracle_eht_chi2(d), 6)
    assert compared >= 6 * 16 * 0.5
    assert np.all(fit.pcur[:, :, 3] == np.float32(0.21)) and np.all(fit.pcur[:, :, 6] == np.float32(0.8))


def test_lockstep_sgra13():
    # the 13-parameter model: its float64 trigonometric invariants (alpha,
    # beta, eta - th, th) are kept across proposals that move other parameters
    d = synthdata.ring_gauss_dataset(nvis=300, ncph=90, seed=26)
    fit = eht_fit(d, nbeta=8, nseq=16, nburn=3, niter=3, seed=88, model="sgra13")
    compared, _ = _lockstep(fit, oracle_eht_chi2(d, imodel=2), 6)
    assert compared >= 6 * 16 * 0.5


def test_lockstep_near_converged_at_plain_fit_size():
    """The regime where decisions are marginal, at the plain fit's data size.

    2000 amplitudes + 600 closure phases (32 tiles, so the step kernel's
    512-visibility LDS ring wraps four times per chi^2), chains started next
    to the truth (chi^2 ~ ndat) with steps small enough that -beta/2 dchi^2 is
    of order one: accept decisions are close calls, and the tie band
    (2e-6 relative chi^2, ~5e-3 nats here) must hold almost none of them."""
    from oracle import cport
    d = synthdata.ring_gauss_dataset(nvis=2000, ncph=600, nsta=8, seed=2)
    nb, ns = 32, 16
    fit = eht_fit(d, nbeta=nb, nseq=ns, nburn=2, niter=2, seed=5, init_chains=False,
                  pstp=1e-4 * (synthdata.EHT_PMAX - synthdata.EHT_PMIN))
    rng = np.random.default_rng(5)
    t = d["truth"].astype(np.float64)
    P = np.clip(t * (1 + 1e-3 * rng.standard_normal((nb, ns, 9))), fit.pmint, fit.pmaxt).astype(np.float32)
    fit.pcur[:] = P
    fit.chi2c[:] = cport.eht_chi2_batch(P.reshape(-1, 9), d).reshape(nb, ns).astype(np.float32)
    fit.flag[:] = 1
    assert np.median(fit.chi2c) < 2 * 2600
    chi2 = lambda Q: cport.eht_chi2_batch(Q, d)
    # chi^2 of the chain states within 2e-5: near the truth every residual
    # d - |V| is only 1-5 % of |V|, so the fp32 model's relative error
    # (~1e-6: Bessel table, hardware sine and cosine) enters each term
    # multiplied by the signal-to-noise ratio (20-100); 1.04e-5 was seen
    compared, counts = _lockstep(fit, chi2, 4, chi2_rtol=2e-5)
    assert counts["decisions"] > 4 * 9 * nb * ns * 0.9
    assert counts["ties"] <= 0.01 * counts["decisions"], counts
    assert compared >= 4 * ns * 0.9
    # the decisions were close calls: both outcomes occur at beta = 1
    assert 0 < fit.n_acpt[:, 0].sum() < 4 * 9 * ns
