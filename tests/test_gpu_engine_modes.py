"""Launch forms and batching give identical chains.

Every chain's arithmetic is fixed by its own Philox counters and its wave's
fixed summation order, so the way chains are packed into launches must not
change a single bit:

* waves per work-group 1, 2, 4 and the persistent work-queue form;
* hipGraph replay of captured iteration cycles vs eager launches;
* several fits of different sizes in one session vs each fit alone;
* the no-recording mode (pout/tout/chi2 NULL) leaves the final state unchanged;
* sequence groups iterated on separate streams vs one stream;
* the autotuner, which times every launch form on real iterations, leaves
  the chains as a plain run does;
* the replica exchange run at the start of the next step launch (the
  default for ring+Gaussian ladders of at most 64 rungs) vs an exchange launch
  after every step launch, over runs cut into segments of every parity and
  fits that end inside a segment.

The generic model.cuh path runs the same EHT fit as the native kernel: its
chi^2 agrees to 2e-6 relative.
"""
import numpy as np
import pytest

from cuda_mcmc_amd import synthdata
from cuda_mcmc_amd.engine import Session
from tests.helpers import eht_fit

pytestmark = pytest.mark.gpu

STATE = ("pcur", "chi2c", "pstp", "n_cnt", "n_acpt", "n_hist", "n_exch", "tcur", "flag", "ptent", "ptentn")


def _fit(nvis, ncph, seed, nb=8, ns=16, nburn=12, niter=8, **kw):
    d = synthdata.ring_gauss_dataset(nvis=nvis, ncph=ncph, seed=seed)
    return eht_fit(d, nb, ns, nburn=nburn, niter=niter, seed=seed + 100, **kw)


def _run(fits, **opt):
    with Session(fits, **opt) as s:
        s.run(max(f.nburn + f.niter for f in fits))
        s.download()
    return fits


def _same(a, b, outputs=True):
    for k in STATE + (("pout", "tout", "chi2") if outputs else ()):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)


def test_launch_forms_bit_identical():
    ref = _run([_fit(700, 200, 41)], waves_per_block=4, persistent=0)[0]
    for opt in (dict(waves_per_block=1, persistent=0), dict(waves_per_block=2, persistent=0),
                dict(waves_per_block=4, persistent=1), dict(waves_per_block=2, persistent=1)):
        got = _run([_fit(700, 200, 41)], **opt)[0]
        _same(ref, got)


def test_graph_replay_bit_identical():
    ref = _run([_fit(500, 150, 42)])[0]
    got = _run([_fit(500, 150, 42)], graph_iters=5)[0]
    _same(ref, got)


def test_batched_fits_match_single_sessions():
    specs = [(300, 90, 51, 4, 8), (1200, 360, 52, 8, 16), (640, 0, 53, 6, 12)]
    alone = [_run([_fit(nv, nc, sd, nb, ns)])[0] for nv, nc, sd, nb, ns in specs]
    for persistent in (0, 1):
        together = _run([_fit(nv, nc, sd, nb, ns) for nv, nc, sd, nb, ns in specs], persistent=persistent)
        for a, b in zip(alone, together):
            _same(a, b)


def test_no_recording_keeps_final_state():
    ref = _run([_fit(400, 120, 61)])[0]
    got = _run([_fit(400, 120, 61, record=False)])[0]
    assert got.pout is None
    _same(ref, got, outputs=False)


def test_generic_model_path_matches_native():
    d = synthdata.ring_gauss_dataset(nvis=500, ncph=150, seed=71)
    nat = eht_fit(d, 4, 8, seed=3)
    gen = eht_fit(d, 4, 8, seed=3)
    with Session([nat]) as s:
        s.download()
    with Session([gen], force_generic=True) as s:
        s.download()
    np.testing.assert_array_equal(nat.pcur, gen.pcur)
    np.testing.assert_allclose(gen.chi2c, nat.chi2c, rtol=2e-6)


def test_autotune_keeps_chains_identical():
    ref = _run([_fit(600, 180, 81)])[0]
    got = _fit(600, 180, 81)
    total = got.nburn + got.niter
    with Session([got]) as s:
        w, p, t = s.autotune(2)   # 6 candidates x 2 real iterations
        assert w in (1, 2, 4) and p in (0, 1) and t > 0
        s.run(total - 12)
        s.download()
    _same(ref, got)


def _quad(n, seed, nb=8, ns=24, nburn=10, niter=10):
    from cuda_mcmc_amd import quadsys
    ds = synthdata.quadsys_dataset(nsys=1, neq=n, n=n, seed=seed)
    return quadsys.quad_fit(ds["Q"][0], ds["b"][0], ds["c"][0], nbeta=nb, nseq=ns, nburn=nburn, niter=niter,
                            seed=seed + 7, record=True)


def test_mixed_kinds_in_one_session():
    # ring+Gaussian fits of both models next to quadratic systems of two sizes:
    # four step launches per iteration (one per kernel, own queue slots in the
    # persistent form) and one exchange launch over ladders of 8, 5 and 12
    # rungs; every fit must come out as it does alone
    def make():
        d = synthdata.ring_gauss_dataset(nvis=400, ncph=120, seed=91)
        return [_fit(700, 200, 92, nb=8, ns=16, nburn=10, niter=10),
                eht_fit(d, 5, 12, nburn=10, niter=10, seed=93, model="sgra13"),
                _quad(8, 94, nb=12), _quad(4, 95)]
    alone = [_run([f])[0] for f in make()]
    for opt in (dict(persistent=0), dict(persistent=1), dict(persistent=1, graph_iters=4)):
        together = _run(make(), **opt)
        for a, b in zip(alone, together):
            _same(a, b)


@pytest.mark.parametrize("split", [7, 15])
def test_resume_in_a_second_session(split):
    # stop after `split` iterations (inside burn-in, or inside sampling with
    # samples already recorded), download, and continue in a new session from
    # the downloaded state: the counter-based streams and the uploaded state
    # give the run of one session, bit for bit
    ref = _run([_fit(500, 150, 71, nburn=10, niter=10)])[0]
    f = _fit(500, 150, 71, nburn=10, niter=10)
    with Session([f]) as s:
        s.run(split)
        s.download()
    f.it0, f.init_chains = split, False
    with Session([f]) as s:
        s.run(f.nburn + f.niter - split)
        s.download()
    _same(ref, f)


@pytest.mark.parametrize("split", [4, 13])
def test_resume_quadratic_system(split):
    # the same for a quadratic system on the row-table step: its step sizes,
    # acceptance counts and histories (staged in LDS and chain-major on the
    # device) go through download and upload in the reference layout
    ref = _run([_quad(8, 77, nb=16, ns=12)])[0]
    f = _quad(8, 77, nb=16, ns=12)
    with Session([f]) as s:
        assert s.fit_info(0)["quad_rows"]
        s.run(split)
        s.download()
    f.it0, f.init_chains = split, False
    with Session([f]) as s:
        s.run(f.nburn + f.niter - split)
        s.download()
    _same(ref, f)


@pytest.mark.parametrize("streams", [2, 3, 8])
def test_sequence_groups_on_streams_bit_identical(streams):
    # sequences cut into groups that iterate on their own streams (with their
    # own iteration clocks) give the chains of one stream: mixed kinds, one
    # fit whose sequence count is not a multiple of the group count, a fit
    # with fewer exchange waves than groups, graph replay and the persistent form
    def make():
        d = synthdata.ring_gauss_dataset(nvis=400, ncph=120, seed=97)
        return [_fit(700, 200, 96, nb=8, ns=40, nburn=10, niter=10),
                eht_fit(d, 5, 13, nburn=10, niter=10, seed=98, model="sgra13"),
                _quad(4, 99, nb=6, ns=20)]
    alone = _run(make())
    for opt in (dict(persistent=0), dict(persistent=1, graph_iters=4)):
        got = _run(make(), streams=streams, **opt)
        for a, b in zip(alone, got):
            _same(a, b)


def _run_segments(fits, segments, **opt):
    with Session(fits, trace_iters=sum(segments), **opt) as s:
        for n in segments:
            s.run(n)
        s.download()
        trace = s.trace
    return fits, trace


@pytest.mark.parametrize("opt", [dict(), dict(persistent=1), dict(graph_iters=3), dict(streams=1),
                                 dict(streams=3, waves_per_block=2)])
def test_exchange_in_step_launch_matches_exchange_launch(opt):
    # segments of 1, 2, 3, 5 and 9 iterations (20 in all).  The fits end at
    # iterations 19 (the last segment's closing launch swaps), 8 (its exchange
    # runs at offset 3 of the segment 6..10, so the closing launch copies the
    # second state set back), 7 (exchange at offset 2: already in the first
    # set), 11 (offset 0 of the last segment, exchange at offset 1), and a
    # 13-parameter fit at 8; a 70-rung ladder keeps its exchange launch of
    # every iteration in the same session.
    def make():
        return [_fit(700, 200, 111, nb=8, ns=16, nburn=7, niter=13),
                _fit(300, 90, 112, nb=5, ns=12, nburn=4, niter=5),
                _fit(500, 0, 113, nb=1, ns=20, nburn=4, niter=4),
                _fit(400, 120, 114, nb=40, ns=3, nburn=6, niter=6),
                _fit(200, 60, 115, nb=70, ns=2, nburn=5, niter=5),
                eht_fit(synthdata.ring_gauss_dataset(nvis=300, ncph=90, seed=116), 6, 10, nburn=5, niter=4,
                        seed=117, model="sgra13")]
    ref, tr_ref = _run_segments(make(), [20], exchange_launch=1)
    got, tr_got = _run_segments(make(), [1, 2, 3, 5, 9], **opt)
    for a, b, ta, tb in zip(ref, got, tr_ref, tr_got):
        _same(a, b)
        for k in ("dec", "chi2", "ex"):
            np.testing.assert_array_equal(ta[k], tb[k], err_msg=k)
