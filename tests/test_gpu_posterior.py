"""Posterior-level check of the whole sampler on the GPU.

A linear least-squares problem (a quadratic system with Q = 0) has the
Gaussian posterior p(x) ~ exp(-chi^2 / 2) with mean (A^T W A)^-1 A^T W y and
covariance (A^T W A)^-1.  The beta = 1 samples of a long MH-RE run (adaptation
during burn-in only, as in the reference) must reproduce both within their
Monte-Carlo error: the mean within 5 standard errors estimated from the
spread of the per-sequence means, the covariance within 10 %.
"""
import numpy as np
import pytest

from cuda_mcmc_amd import quadsys
from cuda_mcmc_amd.engine import Session

pytestmark = pytest.mark.gpu


def test_linear_gaussian_posterior():
    rng = np.random.default_rng(11)
    n, neq, sigma = 4, 6, 0.5
    A = rng.normal(size=(neq, n))
    xtrue = rng.uniform(-1.0, 1.0, size=n)
    y = A @ xtrue + sigma * rng.normal(size=neq)
    # residual y - (x^T 0 x + A x): the quadratic system's c is -y
    fit = quadsys.quad_fit(np.zeros((neq, n, n)), A, -y, pmin=-6.0, pmax=6.0, nbeta=8, nseq=64, nburn=2000,
                           niter=4000, seed=5, sigma=sigma, record=True)
    with Session([fit]) as s:
        s.run(fit.nburn + fit.niter)
        s.download()
    W = np.eye(neq) / sigma ** 2
    cov = np.linalg.inv(A.T @ W @ A)
    mean = cov @ A.T @ W @ y
    samples = fit.pout[:, 0, :, :]                       # [n, nseq, niter] at beta = 1
    per_seq = samples.mean(axis=2)                       # [n, nseq]
    got_mean = per_seq.mean(axis=1)
    se = per_seq.std(axis=1, ddof=1) / np.sqrt(per_seq.shape[1])
    print("mean", got_mean, "expected", mean, "standard error", se)
    assert np.all(np.abs(got_mean - mean) < 5 * se + 1e-3)
    flat = samples.reshape(n, -1)
    got_cov = np.cov(flat)
    print("cov diag", np.diag(got_cov), "expected", np.diag(cov))
    np.testing.assert_allclose(got_cov, cov, atol=0.1 * np.max(np.abs(np.diag(cov))), rtol=0.1)
