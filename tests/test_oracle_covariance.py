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


def _kaiser(ba, bb, f, mu):
    """(b_a + f mu^2)(b_b + f mu^2): the flat-sky Kaiser cross spectrum of populations a, b
    over P(k) (its Legendre coefficients are covariance_coefficient, covariance.c:40-57)."""
    return (ba + f * mu * mu) * (bb + f * mu * mu)


def _mu_space_coefficients(l1, l2, pops, bias, f, n1, n2):
    """The Gaussian covariance of the P-level cross spectra with Poisson shot noise,
    Cov[P_ab, P_cd] ~ (P_ac + d_ac/n_a)(P_bd + d_bd/n_b) + (P_ad + d_ad/n_a)(P_bc + d_bc/n_b)
    (a = pop1, b = pop2, c = pop3, d = pop4; n_a = density1, n_b = density2), projected on
    Legendre polynomials by Gauss-Legendre quadrature in mu.  Returns the cosmic (P^2) and
    mixed (P / n) coefficients with the reference's normalisation 1/4 and 1/8."""
    x, w = np.polynomial.legendre.leggauss(40)
    L1 = np.polynomial.legendre.Legendre.basis(l1)(x)
    L2 = np.polynomial.legendre.Legendre.basis(l2)(x)
    a, b, c, d = pops
    P = lambda p, q: _kaiser(bias[p - 1], bias[q - 1], f, x)  # noqa: E731
    dl = lambda p, q: 1.0 if p == q else 0.0  # noqa: E731
    cosmic = P(a, c) * P(b, d) + P(a, d) * P(b, c)
    mixed = P(a, c) * dl(b, d) / n2 + P(b, d) * dl(a, c) / n1 + P(a, d) * dl(b, c) / n2 + P(b, c) * dl(a, d) / n1
    return np.sum(w * L1 * L2 * cosmic) / 4, np.sum(w * L1 * L2 * mixed) / 8


def _unit_blocks(nl, R, which):
    z = np.zeros((nl * nl, R * R))
    o = np.ones((nl * nl, R * R))
    return (o, z) if which == "pk" else (z, o)


def test_covariance_assemble_populations_mu_space():
    """Populations, both densities and the Kronecker deltas of covariance.c:866-1031 against
    the mu-space form: with unit integrals the mixed and cosmic entries are the angular
    coefficients (signs (-1)^((l2-l1)/2) and (-1)^(l1+l2) of the reference), the Poisson
    entries (2 l1 + 1)/(4 pi) (d13 d24 + (-1)^l2 d14 d23) / (s^2 pix n1 n2 V)."""
    ls, R = [0, 2, 4], 3
    sep = np.array([30.0, 60.0, 90.0])
    bias = np.array([[1.3, 1.9, 2.4, 0.7]])
    f, n1, n2, pix, V = 0.8, 2e-3, 5e-4, 7.0, 3.3e8
    for pops in [(1, 1, 1, 1), (1, 2, 1, 2), (1, 1, 2, 2), (2, 1, 2, 1), (1, 2, 2, 1), (3, 4, 1, 2), (2, 2, 2, 2)]:
        cm_blk = O.covariance_assemble(ls, sep, bias, [f], [1.0], [n1], [V], [pix], *_unit_blocks(3, R, "pk"),
                                       poisson=False, cosmic=False, pops=pops, nbar2=[n2])[0]
        cc_blk = O.covariance_assemble(ls, sep, bias, [f], [1.0], [n1], [V], [pix], *_unit_blocks(3, R, "pk2"),
                                       poisson=False, mixed=False, pops=pops, nbar2=[n2])[0]
        po_blk = O.covariance_assemble(ls, sep, bias, [f], [1.0], [n1], [V], [pix], *_unit_blocks(3, R, "pk"),
                                       mixed=False, cosmic=False, pops=pops, nbar2=[n2])[0]
        a, b, c, d = pops
        for i, l1 in enumerate(ls):
            for j, l2 in enumerate(ls):
                cc, cm = _mu_space_coefficients(l1, l2, pops, bias[0], f, n1, n2)
                s_d = (-1) ** ((l2 - l1) // 2)
                s_s = (-1) ** (l1 + l2)
                assert np.allclose(cm_blk[i, :, j, :], s_d * s_s * cm / V, rtol=1e-11, atol=0), (pops, l1, l2)
                assert np.allclose(cc_blk[i, :, j, :], s_d * cc / V, rtol=1e-11, atol=0), (pops, l1, l2)
                pd = (a == c) * (b == d) + (-1) ** l2 * (a == d) * (b == c)
                exp = np.diag((2 * l1 + 1) / (4 * np.pi) * pd / (sep ** 2 * pix * n1 * n2 * V)) * (l1 == l2)
                assert np.allclose(po_blk[i, :, j, :], exp, rtol=1e-13, atol=1e-300), (pops, l1, l2)


def test_covariance_assemble_default_populations_read_bias1():
    """With the default populations (1, 1, 1, 1) only galaxy_bias1 enters (covariance.c:60-74)."""
    ls, R = [0, 2], 2
    sep = np.array([40.0, 80.0])
    rng = np.random.default_rng(3)
    pk_int, pk2_int = rng.normal(size=(4, R * R)), rng.normal(size=(4, R * R))
    args = ([0.7], [0.9], [1e-3], [2e8], [5.0], pk_int, pk2_int)
    many = O.covariance_assemble(ls, sep, np.array([[1.4, 2.2, 0.3, 3.1]]), *args)
    one = O.covariance_assemble(ls, sep, np.array([1.4]), *args)
    assert np.array_equal(many, one)


def test_covariance_assemble_binned_poisson():
    """The bin-averaged Poisson term (covariance_window): 1 / (((s + p/2)^3 - (s - p/2)^3) / 3)
    in place of 1 / (s^2 p) (covariance.c:931-944); it tends to the latter as p -> 0."""
    ls, sep, pix = [0, 2], np.array([40.0, 80.0]), 6.0
    z = np.zeros((4, 4))
    args = dict(poisson=True, mixed=False, cosmic=False)
    wb = O.covariance_assemble(ls, sep, [1.0], [0.5], [1.0], [1e-3], [1e8], [pix], z, z, window=True, **args)[0]
    nb = O.covariance_assemble(ls, sep, [1.0], [0.5], [1.0], [1e-3], [1e8], [pix], z, z, **args)[0]
    shell = ((sep + pix / 2) ** 3 - (sep - pix / 2) ** 3) / 3
    for i, l in enumerate(ls):
        assert np.allclose(np.diag(wb[i, :, i, :]) / np.diag(nb[i, :, i, :]), sep ** 2 * pix / shell, rtol=1e-13)
    small = O.covariance_assemble(ls, sep, [1.0], [0.5], [1.0], [1e-3], [1e8], [1e-4], z, z, window=True, **args)[0]
    small_nb = O.covariance_assemble(ls, sep, [1.0], [0.5], [1.0], [1e-3], [1e8], [1e-4], z, z, **args)[0]
    assert np.allclose(small, small_nb, rtol=1e-9, atol=0)
