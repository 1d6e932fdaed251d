"""chi^2 gradient, Gauss-Newton and Laplace-covariance kernels vs references.

* gradient vs float64 central differences of the oracle chi^2
  (h = 1e-6 |p|) and vs torch autograd of the float64 torch restatement:
  |g - g_ref| <= 1e-4 |g_ref| per parameter row;
* the gradient kernel's chi^2 equals calc_modchi2's chi2c (same data arrays)
  to 2e-6 relative and the oracle's to 1e-5;
* 2 J^T r of the Gauss-Newton kernel equals the gradient, J^T J is symmetric
  positive semi-definite and its chi^2 equals the gradient kernel's;
* torch.autograd through adjoint.chi2 gives the kernel's gradient;
* Gauss-Newton from a perturbed start on noise-free data returns to the truth;
* the Laplace covariance at the best fit agrees with the covariance of the
  beta = 1 chain samples of an MH-RE run on the same data (marginal standard
  deviations within a factor 1.35).
"""
import numpy as np
import pytest
import torch

from cuda_mcmc_amd import adjoint, mcmc_interf, synthdata
from cuda_mcmc_amd.engine import Session
from cuda_mcmc_amd.models_fp64 import sgra9_vis
from tests import torch_ref
from tests.helpers import eht_fit, oracle_eht_chi2, random_params
from tests.test_oracle_grad import fd_grad

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _data(nvis=400, ncph=120, seed=7, nsta=8):
    return synthdata.ring_gauss_dataset(nvis=nvis, ncph=ncph, nsta=nsta, seed=seed)


def _inner_prior(rng, n):
    # stay away from the prior edges, where the varied parameters are defined
    lo, hi = synthdata.EHT_PMIN.astype(np.float64), synthdata.EHT_PMAX.astype(np.float64)
    span = hi - lo
    return random_params(rng, lo + 0.05 * span, hi - 0.05 * span, n)


def _gpu_grad(d, P):
    data = adjoint.EhtData.from_dataset(d)
    Pt = torch.as_tensor(P, device=DEV).contiguous()
    chi, g = data.chi2_grad(Pt)
    torch.cuda.synchronize()
    return data, chi.cpu().numpy(), g.cpu().numpy()


def test_grad_vs_fd_and_autograd():
    d = _data()
    P = _inner_prior(np.random.default_rng(11), 24)
    data, chi, g = _gpu_grad(d, P)
    f = oracle_eht_chi2(d)
    np.testing.assert_allclose(chi, f(P.astype(np.float64)), rtol=1e-5)
    gfd = fd_grad(f, P)
    Pt = torch.as_tensor(P.astype(np.float64)).requires_grad_(True)
    torch_ref.eht_chi2(Pt, d).sum().backward()
    gat = Pt.grad.numpy()
    for ref, name in ((gfd, "central differences"), (gat, "torch autograd")):
        err = np.linalg.norm(g - ref, axis=1) / np.linalg.norm(ref, axis=1)
        print(name, "max row error", err.max())
        assert np.all(err < 1e-4), (name, err)


def test_grad_chi2_matches_calc_modchi2():
    d = _data(nvis=300, ncph=90, seed=8)
    nb, ns = 4, 8
    P = _inner_prior(np.random.default_rng(12), nb * ns)
    _, chi, _ = _gpu_grad(d, P)
    fit = eht_fit(d, nb, ns)
    pcur = P.reshape(nb, ns, 9).copy()
    chi2c = np.zeros((nb, ns), np.float32)
    flag = np.ones((nb, ns), np.int32)
    ndatm = nb * ns * (fit.ndat + d["nvis"])
    datm = np.zeros(ndatm, np.float32)
    chi2m = np.zeros((nb, ns, fit.ndat), np.float32)
    mcmc_interf.calcmodchi2(fit.coor, fit.dat, fit.std2r, fit.icoor, fit.idat, datm, chi2m, pcur, chi2c, flag,
                            1, fit.coor.size, fit.ndat, fit.icoor.size, fit.idat.size, ndatm, 9, nb, ns)
    np.testing.assert_allclose(chi, chi2c.ravel(), rtol=2e-6)


def test_gn_system_consistent():
    d = _data(nvis=300, ncph=90, seed=9)
    P = _inner_prior(np.random.default_rng(13), 16)
    data, chi, g = _gpu_grad(d, P)
    c2, jtr, jtj = data.gn_system(torch.as_tensor(P, device=DEV))
    c2, jtr, jtj = c2.cpu().numpy(), jtr.cpu().numpy(), jtj.cpu().numpy()
    np.testing.assert_allclose(c2, chi, rtol=2e-6)
    err = np.linalg.norm(2.0 * jtr - g, axis=1) / np.linalg.norm(g, axis=1)
    assert np.all(err < 1e-4), err
    np.testing.assert_allclose(jtj, np.transpose(jtj, (0, 2, 1)), rtol=0, atol=0)
    assert np.all(np.linalg.eigvalsh(jtj) > -1e-6 * np.abs(jtj).max(axis=(1, 2))[:, None])


def test_autograd_function():
    d = _data(nvis=200, ncph=60, seed=10)
    data = adjoint.EhtData.from_dataset(d)
    P = torch.as_tensor(_inner_prior(np.random.default_rng(14), 8), device=DEV).requires_grad_(True)
    chi = adjoint.chi2(P, data)
    w = torch.linspace(0.5, 2.0, 8, device=DEV, dtype=torch.float64)
    (chi * w).sum().backward()
    _, g = data.chi2_grad(P.detach().contiguous())
    torch.testing.assert_close(P.grad, g * w[:, None].float(), rtol=1e-6, atol=0)


def _noise_free(d):
    """The same observation with model-exact amplitudes and closure phases."""
    V = sgra9_vis(d["u"], d["v"], d["truth"])
    ph = np.angle(V)
    cp = d["cpix"]
    cph = np.arctan2(np.sin(ph[cp[:, 0]] + ph[cp[:, 1]] - ph[cp[:, 2]]),
                     np.cos(ph[cp[:, 0]] + ph[cp[:, 1]] - ph[cp[:, 2]]))
    return synthdata.pack_eht(d["u"], d["v"], np.abs(V), d["eamp"], cph, d["ecph"], cp, truth=d["truth"])


def test_gauss_newton_recovers_truth():
    d = _noise_free(_data(nvis=600, ncph=200, seed=21))
    data = adjoint.EhtData.from_dataset(d)
    truth = np.asarray(d["truth"], np.float64)
    rng = np.random.default_rng(5)
    span = synthdata.EHT_PMAX - synthdata.EHT_PMIN
    P0 = (truth + 0.01 * span * rng.standard_normal((8, 9))).astype(np.float32)
    P0 = np.clip(P0, synthdata.EHT_PMIN, synthdata.EHT_PMAX)
    c0 = oracle_eht_chi2(d)(P0.astype(np.float64))
    P, chi = adjoint.gauss_newton(data, P0, synthdata.EHT_PMIN, synthdata.EHT_PMAX, synthdata.EHT_PDESCR,
                                  iters=30)
    print("chi2 start", c0, "end", chi)
    assert np.all(chi < 1e-3 * c0)
    np.testing.assert_allclose(P, np.broadcast_to(truth, P.shape), rtol=0, atol=2e-3 * span.max())


INTERIOR_TRUTH = np.array([2.0, 25.0, 0.6, 0.25, 0.5, 4.0, 0.75, 0.25, 1.0])


def test_laplace_vs_chain_covariance():
    # a source well inside the prior box, so the posterior is not truncated
    d = synthdata.ring_gauss_dataset(nvis=800, ncph=250, nsta=8, seed=31, truth=INTERIOR_TRUTH)
    nb, ns = 8, 64
    fit = eht_fit(d, nb, ns, nburn=1500, niter=2500, seed=3)
    fit.beta[:] = np.geomspace(1.0, 1e-3, nb).astype(np.float32)
    with Session([fit]) as s:
        s.run(fit.nburn + fit.niter)
        s.download()
    best = np.unravel_index(np.argmin(fit.chi2c), fit.chi2c.shape)
    data = adjoint.EhtData.from_dataset(d)
    pbest, _ = adjoint.gauss_newton(data, fit.pcur[best][None, :], synthdata.EHT_PMIN, synthdata.EHT_PMAX,
                                    synthdata.EHT_PDESCR, iters=30)
    cov, vv = adjoint.laplace_covariance(data, pbest, synthdata.EHT_PDESCR)
    sd_lap = np.sqrt(np.diag(cov[0]))
    # beta = 1 samples of the second half of the sampling stage, [seq, sample, param]
    smp = np.transpose(fit.pout[:, 0, :, fit.niter // 2:], (1, 2, 0)).astype(np.float64)
    smp[:, :, 8] = np.angle(np.exp(1j * (smp[:, :, 8] - pbest[0, 8]))) + pbest[0, 8]
    # sequences whose beta = 1 chain sits in the best fit's mode
    near = np.all(np.abs(smp.mean(axis=1)[:, vv] - pbest[0, vv]) < 5.0 * sd_lap, axis=1)
    print("sequences in the mode:", int(near.sum()), "of", ns)
    assert near.sum() >= 8
    s = smp[near][:, :, vv].reshape(-1, len(vv))
    sd_mc = s.std(axis=0)
    ratio = sd_mc / sd_lap
    print("sd chain", sd_mc, "sd laplace", sd_lap, "ratio", ratio)
    assert np.all((ratio > 1 / 1.35) & (ratio < 1.35)), ratio


@pytest.mark.parametrize("cpext,nsta", [(0, 8), (1, 8), (0, 12)])
def test_ring_and_per_closure_gradient_forms_agree(monkeypatch, cpext, nsta):
    # time-ordered data use the single-pass ring form (adjoints accumulated
    # per visibility in collision-free rounds); MHRE_GRAD_RING=0 forces the
    # per-closure form, which recomputes the three visibilities of every
    # closure.  12 stations put one baseline in up to 10 closures of a batch.
    d = synthdata.ring_gauss_dataset(nvis=1400, ncph=600, nsta=nsta, seed=19)
    if cpext:   # closures from their own (u1, v1, u2, v2), appended after the visibility coordinates
        c = d["cpix"]
        uc = np.stack([d["u"][c[:, 0]], d["v"][c[:, 0]], d["u"][c[:, 1]], d["v"][c[:, 1]]], 1)
        d = dict(d, coor=np.concatenate([d["coor"], uc.ravel()]).astype(np.float32), icoor=None, cpext=1)
    P = _inner_prior(np.random.default_rng(13), 40)
    _, chi_r, g_r = _gpu_grad(d, P)
    monkeypatch.setenv("MHRE_GRAD_RING", "0")
    _, chi_c, g_c = _gpu_grad(d, P)
    np.testing.assert_allclose(chi_r, chi_c, rtol=1e-6)
    err = np.linalg.norm(g_r - g_c, axis=1) / np.linalg.norm(g_c, axis=1)
    print("ring vs per-closure max row difference", err.max())
    assert np.all(err < 1e-5)


def test_gradient_kernel_selection_and_fallback():
    # time-ordered data qualify for the ring kernel; the same data in a random
    # visibility order spread closures over the whole set and fall back to the
    # per-closure kernel, with the same chi^2 and gradient
    d = synthdata.ring_gauss_dataset(nvis=1400, ncph=600, nsta=8, seed=23)
    info = adjoint.EhtData.from_dataset(d).info()
    print("time-ordered", info)
    assert info["ring_gradient"] and info["ring_ok"] and 3 <= info["rounds"] <= 16
    nv = d["nvis"]
    perm = np.random.default_rng(3).permutation(nv)
    inv = np.argsort(perm)
    ds = dict(d, coor=np.ascontiguousarray(d["coor"].reshape(nv, 2)[perm].ravel()),
              dat=np.concatenate([d["dat"][:nv][perm], d["dat"][nv:]]),
              std2r=np.concatenate([d["std2r"][:nv][perm], d["std2r"][nv:]]),
              icoor=inv[d["icoor"]].astype(np.int32))
    info_s = adjoint.EhtData.from_dataset(ds).info()
    print("shuffled", info_s)
    assert not info_s["ring_gradient"] and not info_s["ring_ok"]
    P = _inner_prior(np.random.default_rng(29), 24)
    _, chi_a, g_a = _gpu_grad(d, P)
    _, chi_b, g_b = _gpu_grad(ds, P)
    np.testing.assert_allclose(chi_b, chi_a, rtol=1e-5)
    err = np.linalg.norm(g_b - g_a, axis=1) / np.linalg.norm(g_a, axis=1)
    print("shuffled vs time-ordered max row difference", err.max())
    assert np.all(err < 1e-5)


@pytest.mark.parametrize("ring", [True, False])
def test_grad_vs_fd_and_autograd_at_full_size(monkeypatch, ring):
    """The adjoint-gradient workload's own size: 20000 visibilities + 6000
    closure phases, where the kernels sum the invariants' adjoints in fp32 over
    each closure group before adding them to the fp64 totals.  Six points from
    the prior interior and two next to the truth (where the gradient is small
    against its individual terms); both gradient kernels, each within 1e-4 of
    the gradient norm of float64 central differences of the oracle chi^2 (the
    C port) and of torch autograd."""
    if not ring:
        monkeypatch.setenv("MHRE_GRAD_RING", "0")
    d = _data(nvis=20000, ncph=6000, seed=4)
    rng = np.random.default_rng(41)
    span = synthdata.EHT_PMAX - synthdata.EHT_PMIN
    near = d["truth"].astype(np.float64) + 0.01 * span * rng.standard_normal((2, 9))
    near = np.clip(near, synthdata.EHT_PMIN + 0.02 * span, synthdata.EHT_PMAX - 0.02 * span)
    P = np.concatenate([_inner_prior(rng, 6), near.astype(np.float32)])
    data, chi, g = _gpu_grad(d, P)
    assert data.info()["ring_gradient"]   # the data qualify; MHRE_GRAD_RING=0 selects the other kernel
    f = oracle_eht_chi2(d)   # float64 parameters: the FD steps must not be rounded to float32
    # near the truth each amplitude term (dat - |V|)^2 w carries the fp32
    # error of |V| (~1e-7 relative, more where the inner disk cancels much of
    # the outer one) magnified by the signal-to-noise ratio; summed over 26000
    # terms that is ~1e-5 of chi^2
    np.testing.assert_allclose(chi, f(P.astype(np.float64)), rtol=2e-5)
    gfd = fd_grad(f, P)
    Pt = torch.as_tensor(P.astype(np.float64)).requires_grad_(True)
    torch_ref.eht_chi2(Pt, d).sum().backward()
    gat = Pt.grad.numpy()
    for ref, name in ((gfd, "central differences"), (gat, "torch autograd")):
        err = np.linalg.norm(g - ref, axis=1) / np.linalg.norm(ref, axis=1)
        print(name, "row errors", err)
        assert np.all(err < 1e-4), (name, err)
