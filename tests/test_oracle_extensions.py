"""Pins the oracle's spectral combination (SpectralEffectsExtension's linear path) by closed forms.

  * a single bin with I0 = i0 and no detector response gives the monochromatic extinction
    sum_m mu_m p_m (the extension reduces to the nested projector);
  * a detector response r that is the same in every bin shifts the extinction by -ln r;
  * bins whose intensities are zero in every view are skipped and change nothing;
  * two bins of one material, worked out by hand: ln(I0 / (i1 e^{-mu1 p} + i2 e^{-mu2 p}));
  * beam hardening: with a two-bin spectrum the effective attenuation extinction / p falls as p
    grows, from the intensity-weighted mean of mu towards the smaller mu.
"""
import numpy as np
import pytest

import oracle


def test_single_bin_is_monochromatic():
    rng = np.random.default_rng(1)
    p = rng.random((2, 3, 4, 5)) * 4  # two materials, 3 views
    mu = np.array([[0.2, 0.05]])
    inten = np.full((1, 3), 7.0)
    got = oracle.spectral_extinction(p, mu, inten)
    np.testing.assert_allclose(got, 0.2 * p[0] + 0.05 * p[1], rtol=1e-12)


def test_uniform_response_shifts_by_log():
    rng = np.random.default_rng(2)
    p = rng.random((1, 2, 6))
    mu = np.array([[0.3], [0.1], [0.05]])
    inten = rng.random((3, 2)) + 0.5
    a = oracle.spectral_extinction(p, mu, inten)
    b = oracle.spectral_extinction(p, mu, inten, response=[0.8, 0.8, 0.8])
    np.testing.assert_allclose(b - a, -np.log(0.8), rtol=1e-12)


def test_zero_bins_are_skipped():
    rng = np.random.default_rng(3)
    p = rng.random((2, 4, 8))
    mu = np.array([[0.3, 0.1], [5.0, 9.0], [0.1, 0.2]])
    inten = np.array([[1.0] * 4, [0.0] * 4, [2.0] * 4])
    a = oracle.spectral_extinction(p, mu, inten)
    b = oracle.spectral_extinction(p, mu[[0, 2]], inten[[0, 2]])
    np.testing.assert_array_equal(a, b)


def test_two_bins_by_hand():
    p = np.array([[[0.0, 1.0, 2.5]]])  # one material, one view, three rays
    mu = np.array([[0.4], [0.1]])
    inten = np.array([[3.0], [1.0]])
    want = [np.log(4.0 / (3.0 * np.exp(-0.4 * x) + 1.0 * np.exp(-0.1 * x))) for x in (0.0, 1.0, 2.5)]
    np.testing.assert_allclose(oracle.spectral_extinction(p, mu, inten)[0], want, rtol=1e-12)


def test_beam_hardening():
    x = np.linspace(0.01, 40, 200)
    p = x.reshape(1, 1, -1)
    e = oracle.spectral_extinction(p, [[0.5], [0.1]], [[1.0], [1.0]])[0]
    mu_eff = e / x
    assert np.all(np.diff(mu_eff) < 0)
    assert mu_eff[0] == pytest.approx(0.3, rel=1e-2) and mu_eff[-1] < 0.12
