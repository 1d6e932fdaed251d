"""uv reader and closure builder: oracle restatement, product implementation and
independent counting on the reference's example uv files (tests/golden/)."""
import os
from math import comb

import numpy as np
import pytest

from oracle import uvdata as ouv
from cuda_mcmc_amd import uvdata as puv

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FILES = ["000008_1_1_000snd_uvdata.txt", "090018_1_1_120snd_uvdata.txt"]


@pytest.mark.parametrize("name", FILES)
def test_reader_and_closures(name):
    path = os.path.join(GOLD, name)
    o = ouv.read_uvdata(path)
    p = puv.read_uvdata(path)
    for a, b in zip(o[:8], p[:8]):
        np.testing.assert_array_equal(a, b)
    ulam, vlam, _, amp, phase, blin, tsec, sig = p[:8]
    assert len(amp) == 323
    # raw text check of the first row
    first = open(path).readline().split()
    assert abs(ulam[0] - 1e-3 * float(first[1])) < 1e-6 * abs(ulam[0])
    assert abs(phase[0] - np.radians(float(first[8].rstrip(")")))) < 1e-6
    assert abs(amp[0] - float(first[7].rstrip(","))) < 1e-6
    assert tuple(blin[0]) == tuple(int(x) for x in first[4].split("-"))
    oc = ouv.calc_closures(phase, tsec, blin)
    pc = puv.calc_closures(phase, tsec, blin)
    for a, b in zip(oc, pc):
        np.testing.assert_array_equal(a, b)
    # independent count: sum over equal-time spans of C(#antennas, 3)
    cnt = 0
    for i0, i1 in puv.time_spans(tsec):
        cnt += comb(len(np.unique(blin[i0:i1])), 3)
    assert len(pc[0]) == cnt
    # each closure is phi12 + phi13 - phi23 of stations a1 < a2 < a3 at one time
    for (i, j, k) in pc[1][:50]:
        assert tsec[i] == tsec[j] == tsec[k]
        a1, a2 = blin[i]
        assert blin[j][0] == a1 and blin[k][0] == a2 and blin[j][1] == blin[k][1]


def test_snr_closure_sigma_table():
    pherr = np.loadtxt(os.path.join(GOLD, "pherr_vs_snr.txt"))
    assert pherr.size == 3001
    amp = np.array([1.0, 2.0, 3.0, 400.0, 400.0, 400.0])
    eamp = np.array([0.1, 0.1, 0.1, 1.0, 1.0, 1.0])
    cpix = np.array([[0, 1, 2], [3, 4, 5]])
    sig = puv.snr_closure_sigma(amp, eamp, cpix, pherr)
    s = np.array([10.0, 20.0, 30.0])
    snrc = s.prod() / np.sqrt((s[0] * s[1]) ** 2 + (s[0] * s[2]) ** 2 + (s[1] * s[2]) ** 2)
    assert abs(sig[0] - np.radians(pherr[100 * int(snrc)])) < 1e-7
    assert abs(sig[1] - np.sqrt(3) / 400.0) < 1e-7
