"""Pins of the oracle's FFTLog (twofast.c) and I_l^n tables (integrals.c):
Lanczos lgamma and the gamma-ratio kernel against mpmath, the window and q
selection against their closed forms, the transform against high-precision
quadrature, and the tables against the reference's own benchmark fixtures."""
import os

import mpmath as mp
import numpy as np
import pytest

import oracle as O
from tests.synth import batch_spectra

GOLD = os.path.join(os.path.dirname(__file__), "golden")
H = 0.67
COFFE_H0 = 1 / 2997.92458


def test_lanczos_lgamma_vs_mpmath():
    mp.mp.dps = 30
    for z in [0.975 - 3j, 0.525 + 1800j, 0.3 + 100j, 5 + 0.1j, 1.2 - 40j, 0.633 + 7j, -2.3 + 4j]:
        got = O.twofast_lgamma(z)
        ref = complex(mp.loggamma(z))
        # the reflection branch may differ from the principal loggamma by 2 pi i k
        d = got - ref
        d = complex(d.real, (d.imag + np.pi) % (2 * np.pi) - np.pi)
        assert abs(d) <= 2e-13 * max(1.0, abs(ref)), z


@pytest.mark.parametrize("nl", [(0, 0), (0, 2), (0, 4), (1, 1), (1, 3), (2, 0), (2, 2), (3, 1)])
def test_kernel_u_is_mellin_transform_of_jl(nl):
    """u_l(n) alpha^(it - q) with u = 2^(n-1) sqrt(pi) Gamma((1+l+n)/2) / Gamma((2+l-n)/2),
    n = q - 1 - it (the Mellin transform of j_l, DLMF 10.22.43 with j_l = sqrt(pi/2x) J)."""
    mp.mp.dps = 30
    nu, l = nl
    q = O.twofast_selectqnu(l, nu)
    alpha = 1e-5 * H
    for t in (0.0, 0.7, 13.0, 250.0, 1800.0):
        n = mp.mpc(q - 1, -t)
        ref = 2 ** (n - 1) * mp.sqrt(mp.pi) * mp.gamma((1 + l + n) / 2) / mp.gamma((2 + l - n) / 2)
        ref *= mp.power(alpha, mp.mpc(-q, t))
        got = O.twofast_mql(t, q, l, alpha)
        assert abs(got - complex(ref)) <= 1e-10 * abs(complex(ref)), (nl, t)


def test_selectqnu_values():
    # q + nu = 1.94995 for six default tables, 3.26667 for (n, l) = (2, 0), (3, 1)
    for nu, l in [(0, 0), (0, 2), (0, 4), (1, 1), (1, 3), (2, 2)]:
        assert abs(O.twofast_selectqnu(l, nu) + nu - 1.94995) < 1e-14
    for nu, l in [(2, 0), (3, 1)]:
        assert abs(O.twofast_selectqnu(l, nu) + nu - 3.2666666666666666) < 1e-14


def test_window_closed_form():
    xmin, xl, xr, xmax = 1.0, 2.0, 8.0, 10.0
    assert O.twofast_window(5.0, xmin, xmax, xl, xr) == 1.0
    assert O.twofast_window(1.0, xmin, xmax, xl, xr) == 0.0
    assert O.twofast_window(10.0, xmin, xmax, xl, xr) == 0.0
    for v in (1.3, 1.5, 1.9):
        rho = (v - xmin) / (xl - xmin)
        assert abs(O.twofast_window(v, xmin, xmax, xl, xr) - (rho - np.sin(2 * np.pi * rho) / (2 * np.pi))) < 1e-15
    v = 9.4
    rho = (xmax - v) / (xmax - xr)
    assert abs(O.twofast_window(v, xmin, xmax, xl, xr) - (rho - np.sin(2 * np.pi * rho) / (2 * np.pi))) < 1e-15
    with pytest.raises(O.OracleError):
        O.twofast_window(5.0, xmin, xmax, xr, xl)


def _lognormal_pk(k, kc=0.05, s=0.6):
    return 1e4 * np.exp(-(np.log(k / kc)) ** 2 / (2 * s * s)) / k ** 3


@pytest.mark.parametrize("nl,N,tol", [((0, 0), 16384, 5e-9), ((0, 2), 16384, 5e-10), ((1, 1), 16384, 5e-9),
                                      ((2, 0), 16384, 5e-7), ((0, 0), 4096, 1e-7)])
def test_fftlog_vs_high_precision_quadrature(nl, N, tol):
    """I_l^n(r) of a log-normal spectrum against mpmath quadrature of k^2 P j_l(kr)/(kr)^n."""
    mp.mp.dps = 20
    n, l = nl
    kc, s = 0.05, 0.6
    k = np.logspace(-5, 1, 16384)
    tx, ty = O.integral_table(n, l, k, _lognormal_pk(k), N, small_r=False)
    for r in (5.0, 20.0, 60.0, 150.0):
        f = lambda kk: (kk ** 2 * 1e4 * mp.e ** (-(mp.log(kk / kc)) ** 2 / (2 * s * s)) / kk ** 3 *
                        mp.sqrt(mp.pi / (2 * kk * r)) * mp.besselj(l + 0.5, kk * r) / (kk * r) ** n)
        ref = mp.quad(f, [1e-5] + [kc * mp.e ** (j * 0.5) for j in range(-12, 12)] + [10]) / (2 * mp.pi ** 2)
        if n >= l:
            ref *= r ** (n - l)
        got = O.akima(tx, ty, [r])[0]
        assert abs(got - float(ref)) <= tol * abs(float(ref)), (r, got, float(ref))


def test_gaussian_closed_form():
    """P = e^{-k^2 R^2}: I_0^0(r) = e^{-r^2/4R^2} / (8 pi^{3/2} R^3), window-limited accuracy."""
    R = 5.0
    k = np.logspace(-5, 1, 4096)
    r, y = O.twofast_1bessel(k, np.exp(-(k * R) ** 2), 10000, 0, 0)
    exact = np.exp(-r ** 2 / (4 * R ** 2)) / (8 * np.pi ** 1.5 * R ** 3)
    m = (r > 1) & (r < 40)
    assert np.max(np.abs(y[m] - exact[m])) <= 2e-3 * exact.max()


def test_output_grid_and_linearity():
    k = np.logspace(-5, 1, 2048)
    p1 = _lognormal_pk(k)
    p2 = np.exp(-k)
    r, y1 = O.twofast_1bessel(k, p1, 4096, 2, 0)
    _, y2 = O.twofast_1bessel(k, p2, 4096, 2, 0)
    _, y3 = O.twofast_1bessel(k, 2 * p1 - 3 * p2, 4096, 2, 0)
    assert np.allclose(r, (k[-1] / k[0]) ** (np.arange(4096) / 4096), rtol=1e-14, atol=0)
    assert np.max(np.abs(y3 - (2 * y1 - 3 * y2))) <= 1e-10 * np.max(np.abs(y3))


@pytest.mark.parametrize("index,nl", list(enumerate([(0, 0), (0, 2), (0, 4), (1, 1), (1, 3), (2, 0), (2, 2),
                                                     (3, 1)])))
def test_integral_tables_vs_reference_fixtures(index, nl):
    """COFFE's test_integrals (tests/test_coffe.py:196-233): isclose(rtol=1e-4) on
    1.5 < r < 20000 Mpc, PkL_CLASS input, bessel_bins = 10000."""
    n, l = nl
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    k, pk = k * H, pk / H ** 3
    tx, ty = O.integral_table(n, l, k, pk, 10000, small_r=True)
    x, y = np.loadtxt(os.path.join(GOLD, f"benchmark_integral{index}.dat")).T
    r = x / (COFFE_H0 * H)
    got = O.akima(tx, ty, r)
    if n > l:
        got = got * (COFFE_H0 * H) ** (n - l)
    assert np.all(np.isclose(y, got, rtol=1e-4))
    # away from the zero crossings the agreement is much tighter than the test's 1e-4
    big = np.abs(y) > 1e-3 * np.max(np.abs(y))
    assert np.median(np.abs(got[big] - y[big]) / np.abs(y[big])) < 1e-5


def test_small_r_nodes_gaussian_closed_form():
    """The 43 small-r nodes (QAG-GK61, epsrel 1e-5) of P = exp(-k^2 R^2):
    I_0^0(r) = exp(-r^2/4R^2) / (8 pi^(3/2) R^3) at every node; the r = 0 limits
    (1/2pi^2) int k^(2+l-n) P / (2l+1)!! of (1,1), (2,2), (2,0), (3,1) in closed form."""
    R = 2.0
    k = np.logspace(-5, 1.5, 4096)
    pk = np.exp(-(k * R) ** 2)
    tx, ty = O.integral_table(0, 0, k, pk, 4096, small_r=True)
    exact = np.exp(-tx[:43] ** 2 / (4 * R * R)) / (8 * np.pi ** 1.5 * R ** 3)
    assert np.max(np.abs(ty[:43] / exact - 1)) < 1e-7
    m2 = 1 / (8 * np.pi ** 1.5 * R ** 3)          # (1/2pi^2) int k^2 P
    from scipy.special import erf
    # (1/2pi^2) int_kmin^kmax P: the lower limit k = 1e-5 cuts 2e-5 of it
    m0 = np.sqrt(np.pi) / (2 * R) * (erf(k[-1] * R) - erf(k[0] * R)) / (2 * np.pi ** 2)
    for (n, l), want in {(1, 1): m2 / 3, (2, 2): m2 / 15, (2, 0): m0, (3, 1): m0 / 3}.items():
        _, y = O.integral_table(n, l, k, pk, 1024, small_r=True)
        assert abs(y[0] / want - 1) < 1e-7, (n, l)
    # n < l: I_l^n(0) = 0; at small r, j_2(kr) ~ (kr)^2/15 gives I_2^0(r) ~ r^2 (1/2pi^2) int k^4 P / 15
    _, y = O.integral_table(0, 2, k, pk, 1024, small_r=True)
    assert y[0] == 0.0
    m4 = 3 / (16 * np.pi ** 1.5 * R ** 5)          # (1/2pi^2) int k^4 P
    assert abs(y[1] / (1e-6 ** 2 * m4 / 15) - 1) < 1e-6


def test_akima_jvp_against_differences_and_homogeneity():
    """orc_akima_jvp: the derivative of the Akima interpolant in its node values, against
    central differences on random node values (no weight |m_{i+1} - m_i| near zero, so the
    map is smooth there) and, Akima being homogeneous of degree 1, dout = out along dy = y."""
    rng = np.random.default_rng(41)
    x = np.sort(rng.uniform(0, 10, 60))
    y = rng.standard_normal(60)
    xq = rng.uniform(x[0], x[-1], 200)
    out, dout = O.akima_jvp(x, y, y, xq)
    assert np.allclose(out, O.akima(x, y, xq), rtol=0, atol=1e-14)
    assert np.allclose(dout, out, rtol=1e-13, atol=1e-14)
    for trial in range(3):
        dy = rng.standard_normal(60)
        _, an = O.akima_jvp(x, y, dy, xq)
        h = 1e-6
        fd = (O.akima(x, y + h * dy, xq) - O.akima(x, y - h * dy, xq)) / (2 * h)
        assert np.max(np.abs(an - fd)) <= 1e-7 * np.max(np.abs(fd)), trial


def test_tables_jvp_linear_tables_and_differences():
    """orc_tables_at_sep_jvp: along v = P the tangent equals the values (the tables are
    linear in P and Akima homogeneous of degree 1); along a smooth direction it matches
    central differences of orc_tables_at_sep_batch."""
    specs = [(k, p) for k, p, _ in batch_spectra(2, seed=2, lengths=(1024, 2048))]
    tables = [(0, 0), (0, 2), (1, 1), (2, 2)]
    sep = np.linspace(8.0, 400.0, 30)
    nfft = [len(s[0]) for s in specs]
    vals, tang = O.tables_at_sep_jvp(specs, nfft, tables, sep, [s[1] for s in specs])
    assert np.allclose(tang, vals, rtol=1e-12, atol=1e-14 * np.max(np.abs(vals)))
    dirs = [s[1] * np.exp(-0.5 * ((np.log(s[0]) + 3.0) / 1.5) ** 2) for s in specs]
    _, tang = O.tables_at_sep_jvp(specs, nfft, tables, sep, dirs)
    h = 1e-6
    plus = O.tables_at_sep_batch([(s[0], s[1] + h * d) for s, d in zip(specs, dirs)], nfft, tables, sep)
    minus = O.tables_at_sep_batch([(s[0], s[1] - h * d) for s, d in zip(specs, dirs)], nfft, tables, sep)
    fd = (plus - minus) / (2 * h)
    for s in range(2):
        for t in range(len(tables)):
            assert np.max(np.abs(tang[s, t] - fd[s, t])) <= 1e-6 * np.max(np.abs(fd[s, t])), (s, t)
