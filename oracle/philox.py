"""Philox4x32-10 counter-based RNG and the engine's stream layout (oracle side).

Algorithm: J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw, "Parallel
random numbers: as easy as 1, 2, 3", SC'11 -- the Philox4x32 bijection with
R=10 rounds, multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl key increments
0x9E3779B9 / 0xBB67AE85.

Parity: pinned by the three published Random123 known-answer vectors
(``tests/test_oracle_philox.py``).

The reference draws from cuRAND XORWOW states (``src/gpu_mcmc.cu:228-239``,
``:370``, ``:386``, ``:452``, ``:550``, ``:564``); that generator is part of a
closed library and is not reproduced.  Instead the engine defines its own
counter layout, shared bit-for-bit by the device and this oracle:

=====================  ==========================================  ==========
purpose                counter (c0, c1, c2, c3)                    words used
=====================  ==========================================  ==========
MH substep             (chain, iteration, substep, 0)               w0 pick,
                                                                    w1,w2 Box-
                                                                    Muller,
                                                                    w3 accept
MH pick retry r>=1     (chain, iteration, substep, r)               w0
exchange attempt a     (sequence, iteration, a, EXCH)               w0 pair,
                                                                    w1 accept
exchange retry r>=1    (sequence, iteration, a, EXCH + r)           w0
init parameter i       (chain, INIT_ITER, i, INIT)                  w0
=====================  ==========================================  ==========

``chain`` is the reference's through index ``ist = nseq*ibeta + iseq``
(``src/gpu_mcmc.cu:247``); ``sequence`` is ``iseq``; ``iteration`` counts
burn-in iterations first and continues through the sampling stage.  The key is
the 64-bit seed split into (lo, hi) words.
"""
import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint64(0x9E3779B9)
W1 = np.uint64(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)

EXCH = 0x40000000
INIT = 0x80000000
INIT_ITER = 0xFFFFFFFF
MAX_RETRY = 8


def philox4x32(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10.  All arguments broadcast; returns 4 uint32 arrays."""
    c0, c1, c2, c3, k0, k1 = np.broadcast_arrays(
        *[np.asarray(x, dtype=np.uint64) & MASK for x in (c0, c1, c2, c3, k0, k1)])
    c0, c1, c2, c3 = c0.copy(), c1.copy(), c2.copy(), c3.copy()
    k0, k1 = k0.copy(), k1.copy()
    for r in range(10):
        if r:
            k0 = (k0 + W0) & MASK
            k1 = (k1 + W1) & MASK
        p0 = M0 * c0
        p1 = M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
    return tuple(x.astype(np.uint32) for x in (c0, c1, c2, c3))


def key_of(seed):
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return seed & 0xFFFFFFFF, seed >> 32


def u01(w):
    """Uniform in (0, 1]: ((w >> 8) + 1) * 2^-24, exact in fp32 and fp64."""
    w = np.asarray(w, dtype=np.uint64)
    return ((w >> np.uint64(8)) + np.uint64(1)).astype(np.float64) * 2.0 ** -24


def pick(w, n):
    """floor(n * u) with u = (w + 1) * 2^-32 in (0, 1]; equals n only for w = 2^32-1.

    Restates the reference's ``k = (int)(n*curand_uniform())`` with its
    ``while (k >= n)`` redraw (``src/gpu_mcmc.cu:369-372``, ``:549-552``) in
    exact integer arithmetic.
    """
    w = np.asarray(w, dtype=np.uint64)
    return ((np.uint64(n) * (w + np.uint64(1))) >> np.uint64(32)).astype(np.int64)


def normal(w1, w2):
    """Box-Muller (cosine branch) from two (0, 1] uniforms, in float64."""
    u1 = u01(w1)
    u2 = u01(w2)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
