"""Literal restatement of the reference's uv reader and closure builder (oracle side).

read_uvdata    ``src/imgpu.py:1375-1474`` (regex number extraction, 15 numbers
               per data line, kilolambda -> Megalambda, degrees -> radians,
               zero-weight rows keep their slot as zeros).
calc_closures  ``src/imgpu.py:1079-1325`` (combgen / get_time_indices /
               calc_closures: every triangle of every equal-time span of >= 3
               rows, ph1 + ph2 - ph3 with baselines (a1a2), (a1a3), (a2a3)).

Parity: pinned by independent counting on the fixture files
(tests/test_oracle_uvdata.py); the product's vectorised implementation
(cuda_mcmc_amd/uvdata.py) is checked against this one.
"""
import re
from calendar import timegm
from time import strptime

import numpy as np

PK3I = 2 ** 10
_P = re.compile(r'[+-]?\d+\.?\d*[eE][+-]?\d+|[+-]?\d*\.?\d+[eE][+-]?\d+|[+-]?\d+\.?\d*|[+-]?\d*\.?\d+|[+-]?\d+')


def read_uvdata(uv_file):
    lines = open(uv_file).read().splitlines(True)
    n = len(lines)
    amp = np.zeros(n, np.float32); phase = np.zeros(n, np.float32); sig = np.zeros(n, np.float32)
    ulam = np.zeros(n, np.float32); vlam = np.zeros(n, np.float32); wlam = np.zeros(n, np.float32)
    wei = np.zeros(n, np.float32); tsec = np.zeros(n, np.float32)
    chan = np.zeros(n, np.int32); blin = np.zeros((n, 2), np.int32)
    i = 0
    for row in lines:
        nums = _P.findall(row)
        if len(nums) != 15:
            continue
        nums = np.array(nums)
        u, v, w, chan[i], amp[i], phase[i], weig, sig[i] = np.float64(nums[[5, 6, 7, 10, 11, 12, 13, 14]])
        if weig != 0.0:
            wei[i] = weig
            ulam[i] = 1e-3 * u
            vlam[i] = 1e-3 * v
            wlam[i] = 1e-3 * w
            phase[i] = np.radians(phase[i])
            tsec[i] = timegm(strptime(row[:17], '%Y:%j:%H:%M:%S'))
            if row[18:20].isdigit():
                tsec[i] = tsec[i] + 0.01 * np.float32(row[18:20])
            blin[i, :] = int(nums[8]), -int(nums[9])
        i = i + 1
    return (ulam[:i], vlam[:i], wlam[:i], amp[:i], phase[:i], blin[:i], tsec[:i], sig[:i], chan, wei)


def combgen(n, m):
    out = []

    def rec(S, p):
        if p < m:
            start = 0 if p == 0 else S[p - 1] + 1
            for k in range(start, n):
                rec(S + [k], p + 1)
        else:
            out.append(S)
    rec([], 0)
    return np.array(out, int).reshape(-1, m)


def get_time_indices(tsec):
    ntsec = len(tsec)
    itim = np.hstack((0, np.where(np.diff(tsec))[0] + 1, ntsec))
    ditim = np.diff(itim)
    ibeg, iend = [], []
    for i in range(len(ditim)):
        if ditim[i] >= 3:
            ibeg.append(itim[i]); iend.append(itim[i + 1])
    return ibeg, iend


def calc_closures(phase, tsec, blin):
    ibeg, iend = get_time_indices(tsec)
    cphs, cpix, trid, tcs = [], [], [], []
    for i0, i1 in zip(ibeg, iend):
        bl = np.int_(blin[i0:i1, :])
        ant = np.unique(bl)
        itri = combgen(len(ant), 3)
        for j in range(len(itri)):
            a1, a2, a3 = ant[itri[j, :]]
            ib1 = i0 + np.where((bl[:, 0] == a1) & (bl[:, 1] == a2))[0][0]
            ib2 = i0 + np.where((bl[:, 0] == a1) & (bl[:, 1] == a3))[0][0]
            ib3 = i0 + np.where((bl[:, 0] == a2) & (bl[:, 1] == a3))[0][0]
            ph = phase[ib1] + phase[ib2] - phase[ib3]
            cphs.append(np.arctan2(np.sin(ph), np.cos(ph)))
            cpix.append((ib1, ib2, ib3))
            trid.append(a1 + PK3I * (a2 + PK3I * a3))
            tcs.append(tsec[i0])
    return (np.array(cphs, np.float32), np.array(cpix, np.int32).reshape(-1, 3),
            np.array(trid, np.int32), np.array(tcs, float))
