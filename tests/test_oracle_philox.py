"""Philox4x32-10 of the oracle pinned by the published Random123 known-answer vectors."""
import numpy as np

from oracle import philox as px

KAT = [
    # counter (4 words), key (2 words), expected output (Random123 kat_vectors, philox4x32 R=10)
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


def test_known_answer_vectors():
    for ctr, key, want in KAT:
        got = px.philox4x32(*ctr, *key)
        assert tuple(int(g) for g in got) == want


def test_vectorised_matches_scalar():
    c0 = np.arange(100, dtype=np.uint64)
    vec = px.philox4x32(c0, 7, 3, 1, 11, 13)
    for i in (0, 17, 99):
        one = px.philox4x32(i, 7, 3, 1, 11, 13)
        assert all(int(v[i]) == int(o) for v, o in zip(vec, one))


def test_uniform_and_pick_edges():
    assert px.u01(0) == 2.0 ** -24
    assert px.u01(0xFFFFFFFF) == 1.0
    assert np.all(px.u01(np.arange(0, 2 ** 32, 2 ** 20, dtype=np.uint64)) > 0)
    # pick: floor(n u) with u = (w+1) 2^-32, equals n only for w = 2^32-1 (the redraw case)
    for n in (1, 2, 9, 31):
        assert px.pick(0xFFFFFFFF, n) == n
        assert px.pick(0xFFFFFFFE, n) == n - 1
        assert px.pick(0, n) == 0
        w = np.random.default_rng(n).integers(0, 2 ** 32 - 1, 10000, dtype=np.uint64)
        k = px.pick(w, n)
        assert k.min() >= 0 and k.max() < n


def test_normal_moments():
    c = np.arange(200000, dtype=np.uint64)
    w = px.philox4x32(c, 1, 2, 0, 5, 6)
    z = px.normal(w[1], w[2])
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1) < 0.01


def test_redraw_fixture_keys_force_the_redraw():
    """The committed keys make the first word 0xFFFFFFFF at their counters, so
    the pick equals n and the oracle's redraw loop (gpu_mcmc.cu:369-372,
    :549-552) takes the next counter's word; the library's host Philox agrees."""
    import ctypes as C
    import json
    import os
    from cuda_mcmc_amd import _lib
    from oracle.mhre import _pick
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "redraw_keys.json")))
    L = _lib.load()
    for name, n in (("step_pick", 9), ("exchange_pick", 31)):
        e = fx[name]
        c0, c1, c2, c3 = e["counter"]
        w = px.philox4x32(c0, c1, c2, c3, e["key0"], e["key1"])
        assert int(w[0]) == 0xFFFFFFFF, name
        out = (C.c_uint * 4)()
        L.mhre_philox4x32_10((C.c_uint * 4)(c0, c1, c2, c3), (C.c_uint * 2)(e["key0"], e["key1"]), out)
        assert [int(x) for x in out] == [int(x) for x in w]
        assert px.pick(w[0], n) == n
        w1 = px.philox4x32(c0, c1, c2, c3 + 1, e["key0"], e["key1"])[0]
        k = _pick(n, np.array([w[0]], np.uint64),
                  lambda r: np.array([px.philox4x32(c0, c1, c2, c3 + r, e["key0"], e["key1"])[0]], np.uint64))
        assert int(k[0]) == int(px.pick(w1, n)) < n
