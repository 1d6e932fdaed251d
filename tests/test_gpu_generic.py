"""The reference's model.cuh contract and the native quadratic-system kernel,
in lock step with the float64 oracle (same Philox streams):

* Rosenbrock (examples/search_Rosenbrock/model.cuh) and solvsys
  (examples/solvsys/model.cuh) through the generic model() path of their own
  per-project libraries;
* systems of 8 quadratic equations in 8 unknowns on the native kernel;
* Gauss-Newton polish of quadratic systems from near a planted root, and a
  small end-to-end solve (MH-RE, best chain, polish).
Decisions identical except inside the tie band; counters, histories and step
sizes bit-identical on the sequences without a tie.  The quadratic-system
kernel evaluates x^T Q x in fp32 (about 1e-6 of the largest term), so its
tie band is 3e-5 of chi^2 instead of 2e-6, and it draws the Gaussian step in
fp32, so its proposals match the oracle's to a few ulp of the step (parameters
compared to 1e-6 relative or 1e-7 absolute, the prior box being [-5, 5]; a
prior check within 4 ulp of an edge is a tie).
"""
import os

import numpy as np
import pytest

from cuda_mcmc_amd import quadsys, synthdata
from cuda_mcmc_amd.engine import Fit
from oracle.chi2 import generic_chi2
from tests.test_gpu_mhre import _lockstep

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _lib(name):
    p = os.path.join(ROOT, "examples", name, "libmhre.so")
    if not os.path.isfile(p):
        pytest.fail("example library %s not built" % p)
    return p


def test_lockstep_rosenbrock_generic():
    coor = np.zeros(1, np.float32)
    fit = Fit([1, 1], [-5, -5], [5, 5], coor, [0.0], [1.0], nbeta=32, nseq=16, nburn=4, niter=4,
              model="user", npass=1, seed=9, pstp=0.5, adapt_target=0.25,
              beta=np.geomspace(1.0, 1e-3, 32).astype(np.float32))
    f = lambda P: generic_chi2("rosenbrock", P, coor, fit.dat, fit.std2r)
    compared, _ = _lockstep(fit, f, 8, lib=_lib("search_Rosenbrock"))
    assert compared >= 8 * 16 * 0.5


def test_lockstep_solvsys_generic():
    coor = np.array([2, 2, 0.25, 0.5], np.float32)
    fit = Fit([1, 1], [-5, -5], [5, 5], coor, [1.0, 0.5], [1.0, 1.0], nbeta=32, nseq=14, nburn=4, niter=4,
              model="user", npass=1, seed=4)
    f = lambda P: generic_chi2("solvsys", P, coor, fit.dat, fit.std2r)
    compared, _ = _lockstep(fit, f, 8, lib=_lib("solvsys"))
    assert compared >= 8 * 14 * 0.5


@pytest.mark.parametrize("nbeta,nseq", [(16, 32), (1, 100), (70, 16)])
def test_lockstep_quadsys_native(nbeta, nseq):
    # 16 rungs: proposals and exchange in one launch (k_iter_quad); one rung:
    # the same kernel with 64 sequences per wave and no swaps; 70 rungs:
    # lane-per-chain step launch, then the exchange launch.  A 70-rung
    # sequence makes 630 decisions per iteration, so few sequences get through
    # 7 iterations without a single one in the tie band; their decisions are
    # still compared up to their first tie.
    ds = synthdata.quadsys_dataset(nsys=1, seed=3)
    fit = quadsys.quad_fit(ds["Q"][0], ds["b"][0], ds["c"][0], nbeta=nbeta, nseq=nseq, nburn=4, niter=3, seed=21,
                           sigma=1.0)
    f = lambda P: generic_chi2("quadsys", P, fit.coor, fit.dat, fit.std2r)
    # the kernel draws the Gaussian step z in fp32 (hardware log2 and cosine,
    # |dz| ~ 2e-6 at |z| ~ 5): proposals within 1e-5 step sizes of the
    # oracle's float64 ones, prior checks at the edge count as ties
    compared, counts = _lockstep(fit, f, 7, tie_rel=3e-5, edge_ulps=4, pcur_rtol=1e-6, pcur_step=1e-5)
    assert compared >= (7 * nseq * 0.5 if nbeta <= 64 else 1)
    assert counts["ties"] <= 0.01 * counts["decisions"], counts


def test_quad_gauss_newton_from_near_root():
    ds = synthdata.quadsys_dataset(nsys=32, seed=8)
    rng = np.random.default_rng(0)
    X0 = ds["xstar"] + 0.05 * rng.standard_normal(ds["xstar"].shape)
    X, resid = quadsys.gauss_newton(ds["Q"], ds["b"], -ds["c"], X0, iters=25)
    print("residual norms", np.sort(resid)[-5:])
    assert np.all(resid < 1e-5)
    np.testing.assert_allclose(X, ds["xstar"], atol=1e-4)


def test_quad_polish_keeps_the_best_start_per_system():
    # two starts per system, one next to the planted root, one anywhere in
    # the box: the polish keeps the lower residual of the two
    ds = synthdata.quadsys_dataset(nsys=8, seed=12)
    rng = np.random.default_rng(1)
    good = ds["xstar"] + 0.02 * rng.standard_normal(ds["xstar"].shape)
    far = rng.uniform(-5, 5, ds["xstar"].shape)
    X = np.concatenate([far, good]).astype(np.float32)
    sysv = np.concatenate([np.arange(8), np.arange(8)]).astype(np.int32)
    Xb, r = quadsys.polish_best(ds["Q"], ds["b"], -ds["c"], X, sysv, 8, iters=25)
    _, r_good = quadsys.gauss_newton(ds["Q"], ds["b"], -ds["c"], good.astype(np.float32), iters=25)
    _, r_far = quadsys.gauss_newton(ds["Q"], ds["b"], -ds["c"], far.astype(np.float32), iters=25)
    np.testing.assert_array_equal(r, np.minimum(r_good, r_far))
    assert np.all(r < 1e-5)


def test_quad_solve_end_to_end():
    ds = synthdata.quadsys_dataset(nsys=32, seed=9)
    out = quadsys.solve_systems(ds, nbeta=16, nseq=32, nburn=1500, gn_iters=25, seed=100)
    ok = out["resid"] < 1e-5
    print("solved %d of %d; best chi2 median %.3g" % (ok.sum(), ok.size, np.median(out["chi2_best"])))
    assert ok.mean() >= 0.75
