"""CPU restatement of the reference's flat-sky lensing multipoles (test
infrastructure only: tests and the bench's CPU baseline use it as the checker).

  flat_lensing_lensing  functions.c:4020-4133 (the integrand) with signal.c:704-736
                        (2 pi^2 times its integral over x in [0, 1]), where the
                        reference runs GSL QAG to integration_1d_prec = 1e-5; here by
                        brute force, 2048 panels of 32 Gauss-Legendre nodes: the
                        Akima lookups make the integrand piecewise cubic, which leaves
                        adaptive error estimates unreliable below ~1e-9, while the
                        panel rule converges (1024 and 2048 panels agree to ~1e-12)
  flat_density_lensing  functions.c:4136-4257 with signal.c:612-628 (2 pi^2 times the
                        closed form over the half-integer tables)
  half_table            integrals.c:917-960, 1246-1300 with 459-573: the half-integer
                        tables, r^(n - l) I_l^n(r), FFT and small-r nodes
  flat_lensing_lensing_cf, flat_density_lensing_cf
                        functions.c:3515-3558 and 1563-1622: the same terms in the
                        correlation function, over the (1/2, -1/2) table

Every lookup is the oracle's Akima (GSL non-rounded ends), as the reference's splines.
Pinned by the reference's fixtures tests/benchmarks/benchmark_flatsky_{lensing_lensing,
density_lensing}_multipoles{0,2,4}.dat and benchmark_flatsky_lensing_lensing_corrfunc
{0..4}.dat (the reference's own tests hold them to rtol 5e-4), and by the Legendre
projection of the correlation-function terms onto the multipole terms (their tables
differ: the (1/2, -1/2) one against the (1, l) and ((l-2k+3)/2, (l-2k+1)/2) ones), which
tests/test_oracle_lensing.py checks.
"""
from math import comb, factorial, pi, sqrt

import numpy as np

from . import oracle as O


def half_table(n2, l2, k, pk, N, kmin=None, kmax=None, small_r=True):
    """The table (n, l) = (n2/2, l2/2) (integrals.c:459-573 with the half-integer states
    of 186-231, 280-325): the FFT nodes from twofast_1bessel with l and nu
    half-integers, times r^(n - l) when n >= l; with small_r also r = 0 and the 42
    small separations of the integer tables, by scipy quadrature (epsrel 1e-9) of the
    oracle's Akima P(k) against sqrt(pi/2) J_{l+1/2}(kr) / (kr)^(n+1/2), and I(0) =
    sqrt(pi/2) int k^(2+l-n) P / (2^m m!), m = l + 1/2, all over 2 pi^2."""
    from math import factorial as fact

    from scipy.integrate import quad
    from scipy.special import jv
    k, pk = np.asarray(k, dtype=np.float64), np.asarray(pk, dtype=np.float64)
    kmin = float(k[0]) if kmin is None else kmin
    kmax = float(k[-1]) if kmax is None else kmax
    n, l = n2 / 2.0, l2 / 2.0
    x, y = O.twofast_1bessel(k, pk, N, l, n, 1.0, None, kmin, kmax)
    if n2 >= l2:
        y = y * x ** (n - l)
    if not small_r:
        return x, y
    import warnings
    warnings.simplefilter("ignore")   # quad's round-off notes on the Akima (piecewise cubic) P(k)
    m = int(round(l + 0.5))
    P = lambda kk: float(O.akima(k, pk, [kk])[0])   # noqa: E731
    rs = O.integral_table(0, 0, k, pk, 64)[0][:43]     # r = 0 and the 42 small separations
    ys = np.empty(43)
    opts = dict(epsrel=1e-9, epsabs=0.0, limit=1000)
    if n >= l:
        ys[0] = sqrt(pi / 2) * quad(lambda kk: kk ** (2 + l - n) * P(kk), kmin, kmax, **opts)[0] / (2 ** m * fact(m))
    else:
        ys[0] = 0.0
    for i in range(1, 43):
        r = rs[i]
        v = quad(lambda kk: kk * kk * P(kk) * sqrt(pi / 2) * jv(m, kk * r) / (kk * r) ** (n + 0.5), kmin, kmax,
                 **opts)[0]
        ys[i] = v * r ** (n - l) if n >= l else v
    ys /= 2 * pi * pi
    return np.concatenate([rs, x]), np.concatenate([ys, y])


def _lookup(table, v):
    x, y = table
    return float(O.akima(x, y, [v])[0])


def flat_lensing_lensing(table_1l, bg, z_mean, l, sep, s1, s2, Omega_m, H0h, panels=2048):
    """2 pi^2 int_0^1 F(x) dx of functions.c:4020-4133 at one (z_mean, l, sep); table_1l
    is the (1, l) table (x, y); bg a dict of the background samples z, chi, D1; s1, s2
    the magnification biases at z_mean.  Odd l: 0 (signal.c:712)."""
    if l % 2:
        return 0.0
    chi = _lookup((bg["z"], bg["chi"]), z_mean)
    t, w = np.polynomial.legendre.leggauss(32)
    edges = np.linspace(0.0, 1.0, panels + 1)
    a, b = edges[:-1, None], edges[1:, None]
    x = (a + 0.5 * (b - a) * (t + 1)).ravel()
    wx = (0.5 * (b - a) * w).ravel()
    zl = O.akima(bg["chi"], bg["z"], chi * x)           # z_as_chi (background.c:611-617)
    D = O.akima(bg["z"], bg["D1"], zl)
    tv = O.akima(table_1l[0], table_1l[1], x * sep)
    if l > 0:
        tv = tv * x * sep
    val = float(np.sum(wx * tv * D * D * (1 + zl) ** 2 * (x * (1 - x)) ** 2))
    pre = ((2 * l + 1) * factorial(l) / 2 ** l / factorial(l // 2) ** 2 * 9 * Omega_m ** 2 * (2 - 5 * s1) *
           (2 - 5 * s2) * chi ** 3 / 8 / pi * H0h ** 4)
    return 2 * pi * pi * pre * val


def flat_density_lensing(tables_k, bg, z_mean, l, sep, s1, s2, b1, b2, Omega_m, H0h):
    """2 pi^2 times functions.c:4136-4257 at one (z_mean, l, sep); tables_k[k] is the
    half-integer table ((l - 2k + 3)/2, (l - 2k + 1)/2), k <= l/2.  Odd l: 0."""
    if l % 2:
        return 0.0
    s = 0.0
    for k in range(l // 2 + 1):
        s += ((-1) ** k / 2 ** k * comb(l, k) * comb(2 * l - 2 * k, l) * factorial(l // 2 - k) *
              _lookup(tables_k[k], sep))
    D1 = _lookup((bg["z"], bg["D1"]), z_mean)
    s *= (-3 * Omega_m / pi / 8 * (2 * l + 1) * D1 * D1 * (1 + z_mean) * sep *
          ((2 - 5 * s1) * b2 + (2 - 5 * s2) * b1) * 2 ** (1 - l // 2))
    return 2 * pi * pi * s / sqrt(2 * pi) * H0h ** 2


def _cf_table_scale():
    # the (1/2, -1/2) table is rescaled by sqrt(2/pi) 2 pi^2 (integrals.c:956-962)
    return sqrt(2 / pi) * 2 * pi * pi


def flat_lensing_lensing_cf(table_h, bg, z_mean, mu, sep, s1, s2, Omega_m, H0h, panels=2048):
    """The flat-sky lensing-lensing correlation function (functions.c:3515-3558, the
    double-integrated term with flatsky_nonlocal outside MULTIPOLES, times (H0 h)^4 as
    functions.c:3983-3986 returns it) at one (z_mean, mu, sep): the integrand depends on
    x1 alone, so the x2 integral is 1 and x1 takes the brute-force panel rule of
    flat_lensing_lensing; table_h is the (1/2, -1/2) table."""
    chi = _lookup((bg["z"], bg["chi"]), z_mean)
    t, w = np.polynomial.legendre.leggauss(32)
    edges = np.linspace(0.0, 1.0, panels + 1)
    a, b = edges[:-1, None], edges[1:, None]
    x = (a + 0.5 * (b - a) * (t + 1)).ravel()
    wx = (0.5 * (b - a) * w).ravel()
    zl = O.akima(bg["chi"], bg["z"], chi * x)
    D = O.akima(bg["z"], bg["D1"], zl)
    T = _cf_table_scale() * O.akima(table_h[0], table_h[1], x * sep * sqrt(1 - mu * mu))
    val = float(np.sum(wx * T * D * D * (1 + zl) ** 2 * (x * (1 - x)) ** 2))
    return 9 * Omega_m ** 2 * (2 - 5 * s1) * (2 - 5 * s2) * chi ** 3 / 8 / pi * val * H0h ** 4


def flat_density_lensing_cf(table_h, bg, z_mean, mu, sep, s1, s2, b1, b2, Omega_m, H0h):
    """The flat-sky density-lensing correlation function (functions.c:1563-1622, the
    single-integrated term with flatsky_local_nonlocal outside MULTIPOLES; it does not
    depend on the line-of-sight variable, whose integral over [0, 1] is 1)."""
    D1 = _lookup((bg["z"], bg["D1"]), z_mean)
    T = _cf_table_scale() * _lookup(table_h, sep * sqrt(1 - mu * mu))
    return H0h ** 2 * (-3 * Omega_m / pi / 8 * D1 * D1 * (1 + z_mean) * abs(mu) * sep *
                       ((2 - 5 * s1) * b2 + (2 - 5 * s2) * b1) * T)
