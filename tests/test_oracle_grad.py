"""CPU pins of the gradient references used by the GPU gradient tests.

* torch_ref's chi^2 (float64 torch, literal model formulas) equals the
  oracle's chi^2 (float64 NumPy/SciPy restatement) to 1e-12 relative;
* torch_ref's autograd gradient equals float64 central differences of the
  oracle's chi^2 (h = 1e-6 |p|) to 1e-6 of the gradient norm.
"""
import numpy as np
import torch

from cuda_mcmc_amd import synthdata
from tests import torch_ref
from tests.helpers import oracle_eht_chi2, random_params


def _data(seed=3):
    return synthdata.ring_gauss_dataset(nvis=120, ncph=40, nsta=6, seed=seed)


def fd_grad(f, P, rel=1e-6):
    G = np.zeros_like(P, dtype=np.float64)
    for j in range(P.shape[1]):
        h = rel * np.maximum(np.abs(P[:, j]), 1e-3)
        Pp, Pm = P.astype(np.float64).copy(), P.astype(np.float64).copy()
        Pp[:, j] += h
        Pm[:, j] -= h
        G[:, j] = (f(Pp) - f(Pm)) / (2.0 * h)
    return G


def test_torch_ref_chi2_matches_oracle():
    d = _data()
    P = random_params(np.random.default_rng(1), synthdata.EHT_PMIN, synthdata.EHT_PMAX, 6).astype(np.float64)
    ref = oracle_eht_chi2(d)(P)
    got = torch_ref.eht_chi2(torch.as_tensor(P), d).numpy()
    np.testing.assert_allclose(got, ref, rtol=1e-12)


def test_torch_ref_grad_matches_fd():
    d = _data(4)
    P = random_params(np.random.default_rng(2), synthdata.EHT_PMIN, synthdata.EHT_PMAX, 4).astype(np.float64)
    Pt = torch.as_tensor(P).requires_grad_(True)
    torch_ref.eht_chi2(Pt, d).sum().backward()
    g = Pt.grad.numpy()
    gfd = fd_grad(oracle_eht_chi2(d), P)
    err = np.linalg.norm(g - gfd, axis=1) / np.linalg.norm(gfd, axis=1)
    assert np.all(err < 1e-6), err
