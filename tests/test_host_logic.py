"""Host-side logic of the Python drivers (no GPU calls)."""
import os

import numpy as np
import pytest

from cuda_mcmc_amd import imgpu, mcmc_interf as mi, uvdata
from cuda_mcmc_amd.engine import geometric_beta, var_maps

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_beta_ladder_and_var_maps():
    b = geometric_beta(32, 1.0, 1e-4)
    assert b[0] == 1.0 and abs(b[-1] - 1e-4) < 1e-9 and np.all(np.diff(b) < 0)
    ivar, invar = var_maps([1, 0, 2, 1, 0])
    assert list(ivar) == [0, 2, 3] and list(invar) == [0, -1, 1, 2, -1]


def test_mcgpu_defaults_follow_reference():
    m = imgpu.Mcgpu(coor=np.array([2, 2, 0.25, 0.5]), dat=np.zeros(1), pdescr=np.array([1, 1]),
                    pmint=np.array([-10., -10.]), pmaxt=np.array([10., 10.]), nbeta=8, nseq=4, nburn=5, niter=7,
                    seed=5)
    assert m.pstp.shape == (2, 8, 4) and np.allclose(m.pstp, 0.4)
    assert np.all(m.n_cnt == 1) and m.n_hist.shape == (100, 2, 8, 4)
    assert np.all(m.tcur == np.arange(8)[:, None])
    assert np.all(m.pcur == 0.0)        # prior centres
    assert m.pout_4d.shape == (2, 8, 4, 7) and m.chi2_2d.shape == (4, 7)
    assert m.seed == 5 and m.ndatm == 8 * 4 * 1


def test_mcgpu_sgra_packing_from_uv_file():
    path = os.path.join(GOLD, "000008_1_1_000snd_uvdata.txt")
    pd = np.array([1, 1, 1, 1, 1, 1, 1, 1, 2], np.int32)
    lo = np.array([0.5, 20, 0.3, 0, 0, 1.6, 0.5, 0, -np.pi], np.float32)
    hi = np.array([3, 30, 0.9, 0.5, 1, 7.5, 1, 0.5, np.pi], np.float32)
    s = imgpu.Mcgpu_Sgra(uvfile=path, pdescr=pd, pmint=lo, pmaxt=hi, nbeta=4, nseq=2, ecph=0.1)
    assert s.error == ""
    u, v, _, amp, phase, blin, tsec, sig, _, _ = uvdata.read_uvdata(path)
    assert s.nvis == 323 and s.ncph == len(s.cphs) and s.ndat == s.nvis + s.ncph
    np.testing.assert_allclose(s.coor[0::2], 1e-3 * u)
    np.testing.assert_allclose(s.coor[1::2], 1e-3 * v)
    np.testing.assert_array_equal(s.icoor, s.cpix.ravel())
    assert list(s.idat) == [323, 4 * 2 * s.ndat, 0]
    np.testing.assert_allclose(s.std2r[:323], 1 / sig.astype(np.float64) ** 2, rtol=1e-6)
    np.testing.assert_allclose(s.std2r[323:], 100.0, rtol=1e-6)
    assert s.ndatm == 4 * 2 * (s.ndat + s.nvis)
    bad = imgpu.Mcgpu_Sgra(uvfile=path, pdescr=pd, pmint=lo, pmaxt=hi, nbeta=4, nseq=2, eamp=0.0, ecph=0.1)
    assert "zero" in bad.error
    with pytest.raises(ValueError):
        bad.burnin_and_search()
    noecph = imgpu.Mcgpu_Sgra(uvfile=path, pdescr=pd, pmint=lo, pmaxt=hi, nbeta=4, nseq=2)
    assert "ecph" in noecph.error


def test_interface_checks_dtype_shape_contiguity():
    f32 = np.float32
    with pytest.raises(TypeError):
        mi.calcmodel(np.zeros(4), np.zeros(3, np.int32), np.zeros(8, f32), np.zeros((1, 1, 9), f32),
                     1, 4, 2, 3, 8, 9, 1, 1)
    with pytest.raises(ValueError):
        mi.calcmodel(np.zeros(4, f32), np.zeros(3, np.int32), np.zeros(8, f32), np.zeros((1, 9), f32),
                     1, 4, 2, 3, 8, 9, 1, 1)
    with pytest.raises(ValueError):
        a = np.zeros((2, 9), f32)[:, ::2]
        mi.calcmodel(np.zeros(4, f32), np.zeros(3, np.int32), np.zeros(8, f32), a.reshape(1, 2, -1),
                     1, 4, 2, 3, 8, 9, 1, 1)


def test_best_sequence_rows_takes_each_sequences_best_chain():
    from types import SimpleNamespace
    from cuda_mcmc_amd import quadsys
    B, S, N = 3, 5, 2
    chi2c = np.array([[5., 1., 9., 4., 8.],
                      [2., 7., 3., 6., 0.5],
                      [6., 2., 1., 5., 9.]], np.float32)
    pcur = np.arange(B * S * N, dtype=np.float32).reshape(B, S, N)
    f = SimpleNamespace(chi2c=chi2c, pcur=pcur, nseq=S)
    X, c2, sysv = quadsys.best_sequence_rows([f, f], k=3)
    # per sequence best: s0 2.0 (b1), s1 1.0 (b0), s2 1.0 (b2), s3 4.0 (b0), s4 0.5 (b1)
    # three lowest: s4 (0.5), s1 (1.0), s2 (1.0; stable order after s1)
    np.testing.assert_array_equal(c2, [0.5, 1.0, 1.0] * 2)
    np.testing.assert_array_equal(X[:3], [pcur[1, 4], pcur[0, 1], pcur[2, 2]])
    np.testing.assert_array_equal(sysv, [0, 0, 0, 1, 1, 1])
