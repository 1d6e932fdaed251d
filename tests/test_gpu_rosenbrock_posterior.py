"""The 2-D Rosenbrock problem at its stated size through the reference's
model.cuh contract (examples/search_Rosenbrock): 32 temperatures from T = 1 to
T = 1e3, 64 sequences, starts uniform in [-5, 5]^2, step 0.5 adapted toward
a 0.25 acceptance rate, 2000 burn-in + 5000 sampling iterations, Philox seed 0.

chi^2 = f(x, y)^2 with f = 100 (y - x^2)^2 + (1 - x)^2 (dat 0, weight 1), so
the beta = 1 chains sample p(x, y) ~ exp(-f^2 / 2) inside the prior box.  The
first and second moments of the recorded beta = 1 samples must match those of
that density, computed by quadrature on a fine grid, within 5 standard errors
estimated from the spread of the 64 per-sequence moments.
"""
import os

import numpy as np
import pytest

from cuda_mcmc_amd.engine import Fit, Session
from tests.helpers import rosenbrock_moments

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_rosenbrock_posterior_matches_quadrature():
    lib = os.path.join(ROOT, "examples", "search_Rosenbrock", "libmhre.so")
    if not os.path.isfile(lib):
        pytest.fail("example library %s not built" % lib)
    B, S = 32, 64
    fit = Fit([1, 1], [-5, -5], [5, 5], np.zeros(1, np.float32), [0.0], [1.0], nbeta=B, nseq=S, nburn=2000,
              niter=5000, model="user", npass=1, seed=0, pstp=0.5, adapt_target=0.25,
              beta=(10.0 ** (-3.0 * np.arange(B) / (B - 1))).astype(np.float32))
    with Session([fit], lib=lib) as s:
        s.run(fit.nburn + fit.niter)
        s.download()
    xs = fit.pout[0, 0].astype(np.float64)        # [nseq, niter] at beta = 1
    ys = fit.pout[1, 0].astype(np.float64)
    per = {"x": xs.mean(axis=1), "y": ys.mean(axis=1), "xx": (xs * xs).mean(axis=1),
           "xy": (xs * ys).mean(axis=1), "yy": (ys * ys).mean(axis=1)}
    ref = rosenbrock_moments()
    for k, v in per.items():
        se = v.std(ddof=1) / np.sqrt(S)
        print("%-2s sampled %.4f  quadrature %.4f  standard error %.4f" % (k, v.mean(), ref[k], se))
        assert abs(v.mean() - ref[k]) < 5 * se + 2e-3, k
