"""The oracle's MH-RE update rule pinned by known answers and known moments."""
import numpy as np

from oracle import mhre, chi2 as och


def _run(chi2_fn, pdescr, pmin, pmax, beta, nseq, nburn, niter, seed=0, pstp=None, **kw):
    pb = mhre.Problem(chi2_fn, pdescr, pmin, pmax, beta, nseq, seed=seed, **kw)
    st = mhre.State(pb, niter=niter)
    if pstp is not None:
        st.pstp[:] = pstp
    return pb, mhre.run(pb, nburn, niter, st=st)


def test_gaussian_target_moments():
    # chi^2 = x^2 + (y/2)^2 -> beta=1 density N(0,1) x N(0,4); chains at beta=0.25 -> N(0,4) x N(0,16)
    f = lambda P: P[:, 0] ** 2 + (P[:, 1] / 2.0) ** 2
    beta = np.array([1.0, 0.25], np.float32)
    pb, st = _run(f, [1, 1], [-40, -40], [40, 40], beta, nseq=64, nburn=300, niter=1500, seed=1)
    x = st.pout[0, 0].ravel(); y = st.pout[1, 0].ravel()
    assert abs(x.mean()) < 0.08 and abs(y.mean()) < 0.16
    assert abs(x.var() - 1.0) < 0.1 and abs(y.var() - 4.0) < 0.4
    xh = st.pout[0, 1].ravel()
    assert abs(xh.var() - 4.0) < 0.5


def test_correlated_gaussian_covariance():
    S = np.array([[1.0, 0.8], [0.8, 1.0]])
    Si = np.linalg.inv(S)
    f = lambda P: np.einsum("ni,ij,nj->n", P, Si, P)
    pb, st = _run(f, [1, 1], [-10, -10], [10, 10], np.array([1.0], np.float32), nseq=128, nburn=300,
                  niter=1500, seed=2)
    s = np.stack([st.pout[0, 0].ravel(), st.pout[1, 0].ravel()])
    np.testing.assert_allclose(np.cov(s), S, atol=0.1)


def test_rosenbrock_minimum():
    # the search_Rosenbrock example: minimum 0 at (1, 1)
    f = lambda P: och.generic_chi2("rosenbrock", P, None, [0.0], [1.0])
    beta = (10.0 ** (-3.0 * np.arange(8) / 7)).astype(np.float32)
    pb, st = _run(f, [1, 1], [-5, -5], [5, 5], beta, nseq=16, nburn=300, niter=300, seed=0, pstp=0.5,
                  adapt_target=0.25)
    c2 = st.chi2.ravel()
    im = c2.argmin()
    po = st.pout[:, 0].reshape(2, -1)
    assert c2[im] < 1e-4
    np.testing.assert_allclose(po[:, im], [1.0, 1.0], atol=0.05)


def test_solvsys_roots():
    # x^2 + y^2 = 1, -0.25 x + y = 0.5: roots ((-4 -+ sqrt(832))/34, sqrt(1 - x^2))
    # sigma = 0.01 on both equations so the cold chain concentrates at the roots
    f = lambda P: och.generic_chi2("solvsys", P, [2.0, 2.0, 0.25], [1.0, 0.5], [1e4, 1e4])
    beta = (10.0 ** (-6.0 * np.arange(12) / 11)).astype(np.float32)
    pb, st = _run(f, [1, 1], [-5, -5], [5, 5], beta, nseq=16, nburn=400, niter=400, seed=3)
    roots = [((-4 - np.sqrt(832)) / 34, 0.0), ((-4 + np.sqrt(832)) / 34, 0.0)]
    roots = [(x, np.sqrt(1 - x * x)) for x, _ in roots]
    po = st.pout[:, 0].reshape(2, -1)
    good = st.chi2.ravel() < 4.0
    assert good.sum() > 1000
    d = np.min([np.hypot(po[0, good] - x, po[1, good] - y) for x, y in roots], axis=0)
    assert np.all(d < 0.05)
    near = [np.any(np.hypot(po[0, good] - x, po[1, good] - y) < 0.02) for x, y in roots]
    assert all(near), "both roots are found"


def test_adaptation_rule_on_flat_target():
    # chi^2 == 0: every in-prior proposal is accepted (log u <= 0), so each step of
    # parameter k multiplies pstp[k] by 1.01 (rate 1 >= 0.30), rounding to float32 each time
    f = lambda P: np.zeros(P.shape[0])
    pb = mhre.Problem(f, [1, 1, 0, 1], [-1e3] * 4, [1e3] * 4, np.array([1.0], np.float32), 3, seed=9)
    st = mhre.State(pb, niter=0)
    mhre.init_pcur(pb, st)
    st.pcur[:] = 0.0          # far from the prior edges with steps ~1e-3
    st.pstp[:] = 1e-3
    p0 = st.pstp.copy()
    counts = np.zeros((3, 1, 3), int)
    for it in range(20):
        tr = {}
        mhre.iteration(pb, st, it, True, trace=tr)
        for e in tr["mh"]:
            for c in range(3):
                counts[e["k"][c], 0, c] += 1
    for k in range(3):
        for c in range(3):
            v = p0[k, 0, c]
            for _ in range(counts[k, 0, c]):
                v = np.float32(np.float64(v) * 1.01)
            assert st.pstp[k, 0, c] == v
    assert np.all(st.n_cnt == 1 + counts)
    assert np.all(st.n_acpt == counts)


def test_exchange_always_swaps_equal_chi2():
    # equal chi^2 -> x = 0 -> every attempt swaps; B attempts per sequence
    f = lambda P: np.ones(P.shape[0])
    beta = np.array([1.0, 0.5, 0.25, 0.1], np.float32)
    pb = mhre.Problem(f, [1], [-1], [1], beta, 5, seed=4)
    st = mhre.State(pb, niter=1)
    mhre.init_pcur(pb, st)
    st.chi2c[:] = 3.0
    before = st.tcur.copy()
    tr = []
    mhre.exchange(pb, st, 0, False, itr_out=0, trace=tr)
    assert tr[0]["accept"].all()
    # replay the recorded pair picks on the labels
    lab = before.copy()
    for a in range(4):
        for s in range(5):
            k = tr[0]["k"][a][s]
            lab[[k, k + 1], s] = lab[[k + 1, k], s]
    np.testing.assert_array_equal(st.tcur, lab)
    assert st.n_exch.sum() == 4 * 5
    np.testing.assert_array_equal(st.tout[:, :, 0], st.tcur)


def test_rosenbrock_posterior_matches_quadrature():
    # the Rosenbrock problem of the GPU test (32 temperatures T = 1 .. 1e3,
    # 64 sequences, step 0.5 adapted toward 0.25, Philox seed 0), with 2000
    # sampling iterations instead of 5000 to keep the suite short: the
    # beta = 1 samples' moments against quadrature of exp(-f^2 / 2)
    from tests.helpers import rosenbrock_moments
    f = lambda P: och.generic_chi2("rosenbrock", P, None, [0.0], [1.0])
    B = 32
    beta = (10.0 ** (-3.0 * np.arange(B) / (B - 1))).astype(np.float32)
    pb, st = _run(f, [1, 1], [-5, -5], [5, 5], beta, nseq=64, nburn=2000, niter=2000, seed=0, pstp=0.5,
                  adapt_target=0.25)
    xs = st.pout[0, 0].astype(np.float64)
    ys = st.pout[1, 0].astype(np.float64)
    per = {"x": xs.mean(axis=1), "y": ys.mean(axis=1), "xx": (xs * xs).mean(axis=1),
           "xy": (xs * ys).mean(axis=1), "yy": (ys * ys).mean(axis=1)}
    ref = rosenbrock_moments()
    for k, v in per.items():
        se = v.std(ddof=1) / np.sqrt(v.size)
        assert abs(v.mean() - ref[k]) < 5 * se + 2e-3, (k, v.mean(), ref[k], se)
