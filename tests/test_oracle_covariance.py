"""Pins of the oracle's 2D FFTLog (twobessel.c / utils.c) and covariance pieces
(covariance.c): the Gaussian closed form of int dk k^2 e^{-k^2 R^2} j_0(k r1) j_0(k r2),
g_l against mpmath, coefficients and the four-Legendre integral."""
import mpmath as mp
import numpy as np

import oracle as O


def test_g_l_vs_mpmath():
    mp.mp.dps = 30
    for l, nu in [(0, 1.01), (2, 1.01), (4, 1.01)]:
        eta = np.array([0.0, 0.3, 5.0, 120.0])
        got = O.g_l(l, nu, eta)
        for e, gv in zip(eta, got):
            z = mp.mpc(nu, e)
            ref = 2 ** z * mp.gamma((l + z) / 2) / mp.gamma((3 + l - z) / 2)
            assert abs(gv - complex(ref)) <= 1e-11 * abs(complex(ref))


def test_two_sph_bessel_gaussian_closed_form():
    R = 8.0
    k = np.logspace(-4, 1, 800)
    pk = np.exp(-(k * R) ** 2)
    sep = np.array([20.0, 30.0, 50.0, 80.0])
    res = O.covariance_fftlog_integral(k, pk, 1.0, 0, 0, 1024, sep)   # [n, m] = value at (sep_m, sep_n)
    r1, r2 = sep[None, :], sep[:, None]
    exact = np.sqrt(np.pi) * (np.exp(-(r1 - r2) ** 2 / (4 * R * R)) - np.exp(-(r1 + r2) ** 2 / (4 * R * R))) / (
        4 * R * r1 * r2)
    assert np.max(np.abs(res - exact)) <= 1e-3 * np.max(np.abs(exact))
    # symmetric in (r1, r2) for l1 = l2, up to the asymmetry of the bicubic cross
    # derivative (d/dy of the x-spline derivatives)
    assert np.max(np.abs(res - res.T)) <= 1e-7 * np.max(np.abs(res))


def test_two_sph_bessel_rejects_odd_sizes():
    k = np.logspace(-3, 0, 9)
    try:
        O.two_sph_bessel(k, k, np.zeros((9, 9)), 0, 0)
    except O.OracleError:
        return
    raise AssertionError("odd N accepted")


def test_covariance_coefficients():
    b1, b2, f = 1.3, 1.7, 0.8
    assert abs(O.lib().orc_covariance_coefficient(b1, b2, f, 0) - (b1 * b2 + (b1 + b2) * f / 3 + f * f / 5)) < 1e-15
    assert abs(O.lib().orc_covariance_coefficient(b1, b2, f, 2) - (2 * (b1 + b2) * f / 3 + 4 * f * f / 7)) < 1e-15
    assert abs(O.lib().orc_covariance_coefficient(b1, b2, f, 4) - 8 * f * f / 35) < 1e-15
