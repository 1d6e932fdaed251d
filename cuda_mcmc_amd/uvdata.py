"""EHT uv-data front end: the reference's text reader and closure builder.

read_uvdata(path)   parses the 15-number-per-line uv text format of
                    ``src/imgpu.py:1375-1474`` (time, u, v, w in kilolambda,
                    baseline ``aa-bb``, channel, (amp, phase in degrees),
                    weight, sigma); u, v, w are scaled to Megalambda and the
                    phase to radians.  Rows with zero weight stay as zeros
                    but keep their slot, as in the reference (``i += 1``
                    outside the weight test, ``imgpu.py:1431-1445``).
calc_closures(...)  all closure triangles of every equal-time span with at
                    least three baselines (``src/imgpu.py:1184-1325``): for
                    antennas a1 < a2 < a3 the baselines (a1a2), (a1a3),
                    (a2a3) and the phase combination ph1 + ph2 - ph3 wrapped
                    with atan2(sin, cos).  Triangles whose three baselines
                    are not all present are skipped (the reference raises
                    IndexError there).
snr_closure_sigma   closure-phase sigma from visibility SNRs (the rule of
                    ``examples/sgra_model_fit/model_obsfit.py:265-283``).
"""
import calendar
import itertools
import re
import time

import numpy as np

PK3I = 2 ** 10

_NUM = re.compile(r"[+-]?\d+\.?\d*[eE][+-]?\d+|[+-]?\d*\.?\d+[eE][+-]?\d+|[+-]?\d+\.?\d*|[+-]?\d*\.?\d+|[+-]?\d+")


def read_uvdata(path):
    """Returns ulam, vlam, wlam, amp, phase, blin, tsec, sig, chan, wei (float32/int32)."""
    rows = []
    with open(path, "r") as fh:
        for row in fh:
            nums = _NUM.findall(row)
            if len(nums) != 15:
                continue
            rows.append((row, nums))
    n = len(rows)
    f32 = np.float32
    amp = np.zeros(n, f32)
    phase = np.zeros(n, f32)
    sig = np.zeros(n, f32)
    ulam = np.zeros(n, f32)
    vlam = np.zeros(n, f32)
    wlam = np.zeros(n, f32)
    wei = np.zeros(n, f32)
    tsec = np.zeros(n, f32)
    chan = np.zeros(n, np.int32)
    blin = np.zeros((n, 2), np.int32)
    for i, (row, nums) in enumerate(rows):
        u, v, w, ch, a, p, weig, s = [float(nums[k]) for k in (5, 6, 7, 10, 11, 12, 13, 14)]
        chan[i] = int(ch)
        amp[i] = a
        phase[i] = p
        sig[i] = s
        if weig != 0.0:
            wei[i] = weig
            ulam[i] = 1e-3 * u
            vlam[i] = 1e-3 * v
            wlam[i] = 1e-3 * w
            phase[i] = np.radians(phase[i])
            t = calendar.timegm(time.strptime(row[:17], "%Y:%j:%H:%M:%S"))
            if row[18:20].isdigit():
                t = t + 0.01 * np.float32(row[18:20])
            tsec[i] = t
            blin[i] = int(nums[8]), -int(nums[9])
    return ulam, vlam, wlam, amp, phase, blin, tsec, sig, chan, wei


def time_spans(tsec):
    """[start, end) of every equal-time span with at least three rows (imgpu.py:1184-1225)."""
    tsec = np.asarray(tsec)
    n = len(tsec)
    starts = np.concatenate([[0], np.nonzero(np.diff(tsec))[0] + 1, [n]])
    return [(int(a), int(b)) for a, b in zip(starts[:-1], starts[1:]) if b - a >= 3]


def calc_closures(phase, tsec, blin):
    """Returns cphs (float32), cpix [n,3] (int32), trid (int32), tcphsec (float64)."""
    phase = np.asarray(phase, np.float32)
    blin = np.asarray(blin)
    cphs, cpix, trid, tcs = [], [], [], []
    if len(phase) >= 3:
        for i0, i1 in time_spans(tsec):
            bl = blin[i0:i1].astype(int)
            first = {}
            for r in range(i1 - i0):
                first.setdefault((bl[r, 0], bl[r, 1]), i0 + r)
            ant = np.unique(bl)
            for a1, a2, a3 in itertools.combinations(ant, 3):
                try:
                    ib = (first[(a1, a2)], first[(a1, a3)], first[(a2, a3)])
                except KeyError:
                    continue
                ph = phase[ib[0]] + phase[ib[1]] - phase[ib[2]]
                cphs.append(np.arctan2(np.sin(ph), np.cos(ph)))
                cpix.append(ib)
                trid.append(a1 + PK3I * (a2 + PK3I * a3))
                tcs.append(tsec[i0])
    return (np.array(cphs, np.float32), np.array(cpix, np.int32).reshape(-1, 3),
            np.array(trid, np.int32), np.array(tcs, np.float64))


def snr_closure_sigma(amp, eamp, cpix, pherr_deg=None):
    """Closure-phase 1-sigma (radians) from the three visibility SNRs.

    SNRc = s1 s2 s3 / sqrt((s1 s2)^2 + (s1 s3)^2 + (s2 s3)^2); above 30 the
    sigma is 1/SNRc, otherwise pherr_deg[100*int(SNRc)] degrees from the
    phase-error table (0..30 in 0.01 steps) when given, else 1/SNRc.
    """
    snr = np.asarray(amp, np.float64) / np.asarray(eamp, np.float64)
    i1, i2, i3 = cpix[:, 0], cpix[:, 1], cpix[:, 2]
    s1, s2, s3 = snr[i1], snr[i2], snr[i3]
    snrc = s1 * s2 * s3 / np.sqrt((s1 * s2) ** 2 + (s1 * s3) ** 2 + (s2 * s3) ** 2)
    sig = 1.0 / snrc
    if pherr_deg is not None:
        low = snrc <= 30.0
        sig[low] = np.radians(np.asarray(pherr_deg)[100 * np.int_(snrc[low])])
    return sig.astype(np.float32)
