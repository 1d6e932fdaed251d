"""The C / OpenMP oracle port against the NumPy oracle (bit-exact state, chi^2 to 1e-13)."""
import numpy as np

from cuda_mcmc_amd import synthdata
from oracle import cport, chi2, mhre
from tests.helpers import eht_fit, oracle_eht_chi2

STATE = ('pcur', 'chi2c', 'pstp', 'n_cnt', 'n_acpt', 'n_hist', 'n_exch', 'tcur', 'flag', 'ptent', 'ptentn')


def test_chi2_batch():
    d = synthdata.ring_gauss_dataset(nvis=300, ncph=90, seed=4)
    rng = np.random.default_rng(0)
    P = (synthdata.EHT_PMIN + (synthdata.EHT_PMAX - synthdata.EHT_PMIN) * rng.uniform(size=(20, 9))).astype(np.float32)
    a = cport.eht_chi2_batch(P, d, nthreads=2)
    b = chi2.eht_chi2(P, d['coor'], d['dat'], d['std2r'], d['icoor'], d['nvis'], d['ncph'])
    np.testing.assert_allclose(a, b, rtol=1e-13)


def test_iterations_bit_exact():
    d = synthdata.ring_gauss_dataset(nvis=300, ncph=90, seed=4)
    fit = eht_fit(d, 4, 8, nburn=2, niter=2, seed=3, nadj=3)
    pb = mhre.Problem(oracle_eht_chi2(d), fit.pdescr, fit.pmint, fit.pmaxt, fit.beta, fit.nseq, nadj=3, seed=3)
    st = mhre.State(pb, niter=2)
    mhre.init_pcur(pb, st)
    for k in STATE:
        setattr(fit, k, getattr(st, k).copy())
    for it in range(4):
        cport.eht_iteration(fit, d, it, nthreads=2)
        mhre.iteration(pb, st, it, it < 2, itr_out=None if it < 2 else it - 2)
    for k in STATE + ('pout', 'tout', 'chi2'):
        np.testing.assert_array_equal(getattr(fit, k), getattr(st, k), err_msg=k)
