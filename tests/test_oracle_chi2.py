"""Closures and chi^2 terms of the oracle against hand-computed cases."""
import numpy as np

from oracle.chi2 import wrap, eht_datm, eht_chi2_terms


def test_wrap():
    x = np.array([0.0, 3.0, 3.2, -3.2, 7.0, -7.0, np.pi / 2])
    np.testing.assert_allclose(wrap(x), np.angle(np.exp(1j * x)), atol=1e-15)
    assert np.all(np.abs(wrap(np.linspace(-50, 50, 1001))) <= np.pi)


def test_three_station_closure_by_hand():
    # one time, stations 1,2,3: baselines 12, 13, 23; closure = ph12 + ph13 - ph23
    prm = np.array([1.3, 24.0, 0.5, 0.3, 0.2, 3.0, 0.8, 0.3, 0.7])
    u = np.array([1.0, 2.5, 1.5])
    v = np.array([0.5, -1.0, -1.5])
    coor = np.empty(6)
    coor[0::2], coor[1::2] = u, v
    amp, cph, pha = eht_datm(prm, coor, [0, 1, 2], 3, 1)
    want = wrap(pha[0] + pha[1] - pha[2])
    assert abs(cph[0] - want) < 1e-15
    # external form: the same closure from (u1,v1),(u2,v2) and -(u1+u2, v1+v2)
    c2 = np.concatenate([coor, [u[0], v[0], u[1], v[1]]])
    _, cph_ext, _ = eht_datm(prm, c2, None, 3, 1, cpext=1)
    from oracle.models import sgra_model9
    p3 = np.angle(sgra_model9(-(u[0] + u[1]), -(v[0] + v[1]), prm))
    assert abs(cph_ext[0] - wrap(pha[0] + pha[1] - p3)) < 1e-12


def test_chi2_terms_weighting_and_wrap():
    prm = np.array([1.0, 25.0, 0.5, 0.2, 0.5, 3.0, 0.8, 0.2, 0.3])
    coor = np.array([1.0, 0.0, 0.0, 1.0, 1.0, 1.0])
    amp, cph, _ = eht_datm(prm, coor, [0, 1, 2], 3, 1)
    dat = np.concatenate([amp + [0.1, -0.2, 0.0], [cph[0] + 2 * np.pi - 0.05]])
    w = np.array([4.0, 1.0, 9.0, 100.0])
    t, _, _ = eht_chi2_terms(prm, coor, dat, w, [0, 1, 2], 3, 1)
    np.testing.assert_allclose(t, [0.04, 0.04, 0.0, 0.25], atol=1e-12)
