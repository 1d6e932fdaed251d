"""Batched sessions against the oracle: uneven fits in one session, each fit
held to the reference's per-fit update rule (src/gpu_mcmc.cu:361-626).

The fits differ in data size (600 to 20000 visibilities with closure phases
at 30 % of that), station count, ladder (16, 20, 24, 33 and 48 rungs; 16, 20
and 24 share an exchange wave between sequences, 33 and 48 take a wave each)
and number of sequences (8 to 128).  Every iteration each fit's oracle restarts from the
state the device downloaded and runs the same iteration; picks must be
identical, accept and swap decisions identical outside the tie band, and the
state of every sequence without a near-tie equal.  Both launch forms of the
step kernel (grid and persistent work queue) are checked.  The oracle's chi^2
is the float64 C port (pinned against the NumPy oracle by test_oracle_c).
"""
import numpy as np
import pytest

from cuda_mcmc_amd import synthdata
from oracle import cport
from tests.helpers import eht_fit
from tests.test_gpu_mhre import _lockstep_batch

pytestmark = pytest.mark.gpu

#        nvis  ncph  nsta  nbeta nseq  seed
FITS = [(5000, 1500, 8, 16, 64, 301),
        (1200, 360, 6, 33, 64, 302),
        (600, 180, 5, 48, 80, 303),
        (900, 270, 10, 20, 128, 304),
        (20000, 6000, 10, 24, 8, 305)]


def _batch():
    fits, fns = [], []
    for nvis, ncph, nsta, nb, ns, seed in FITS:
        d = synthdata.ring_gauss_dataset(nvis=nvis, ncph=ncph, nsta=nsta, seed=seed)
        fits.append(eht_fit(d, nbeta=nb, nseq=ns, nburn=2, niter=1, seed=seed))
        fns.append(lambda P, d=d: cport.eht_chi2_batch(P, d))
    return fits, fns


@pytest.mark.parametrize("persistent", [0, 1])
def test_batched_session_matches_oracle_per_fit(persistent):
    fits, fns = _batch()
    counts = {}
    compared = _lockstep_batch(fits, fns, 3, counts=counts, persistent=persistent)
    nseq_total = sum(f.nseq for f in fits)
    print("sequence states compared", compared, "of", 3 * nseq_total, counts)
    # every decision outside the tie band was compared; a sequence's state is
    # compared only when none of its (nprm + 1) nbeta decisions of the
    # iteration fell in the band, which at these chi^2 (1e5-1e7, random
    # starts) leaves out most long-ladder sequences
    assert counts["ties"] <= 0.02 * counts["decisions"], counts
    assert compared >= 0.1 * 3 * nseq_total, (compared, nseq_total)
    # every fit advanced through burn-in into its sampling stage
    for f in fits:
        assert np.all(f.n_cnt >= 1) and f.tout.shape[-1] == 1
