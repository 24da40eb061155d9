"""Pins of the oracle's primitives: FFTs, splines, quadrature, special functions.

Each check uses an independent implementation (numpy/scipy/mpmath), a closed form or
a brute-force sum, chosen so that a wrong sign, index or dropped term fails."""
import mpmath as mp
import numpy as np
import pytest
import scipy.interpolate as si

import oracle as O

RNG = np.random.default_rng(12345)


def _ld_dft(x, sign=-1):
    """brute-force DFT: exact index reduction (k j mod n), roots of unity from mpmath
    at 30 digits, long-double accumulation."""
    n = len(x)
    mp.mp.dps = 30
    roots_c = np.array([mp.cos(2 * mp.pi * m / n) for m in range(n)], dtype=np.longdouble)
    roots_s = np.array([sign * mp.sin(2 * mp.pi * m / n) for m in range(n)], dtype=np.longdouble)
    k = np.arange(n)
    idx = np.outer(k, k) % n
    c, s = roots_c[idx], roots_s[idx]
    xr, xi = x.real.astype(np.longdouble), x.imag.astype(np.longdouble)
    re = (c * xr - s * xi).sum(axis=1)
    im = (s * xr + c * xi).sum(axis=1)
    return (re + 1j * im).astype(np.complex128)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 8, 12, 30, 64, 97, 101, 210])
def test_fft_bruteforce(n):
    x = RNG.standard_normal(n) + 1j * RNG.standard_normal(n)
    ref = _ld_dft(x)
    assert np.max(np.abs(O.fft(x) - ref)) <= 1e-14 * max(1.0, np.max(np.abs(ref)))
    refi = _ld_dft(x, +1)
    assert np.max(np.abs(O.fft(x, +1) - refi)) <= 1e-14 * max(1.0, np.max(np.abs(refi)))


@pytest.mark.parametrize("n", [1000, 1024, 10000, 16384, 8000, 2 * 3 * 5 * 7 * 11])
def test_fft_numpy(n):
    x = RNG.standard_normal(n) + 1j * RNG.standard_normal(n)
    ref = np.fft.fft(x)
    assert np.max(np.abs(O.fft(x) - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("n", [4, 6, 10, 64, 202, 1024, 4096, 8000, 10000, 16384])
def test_rfft_c2r_fftw_semantics(n):
    x = RNG.standard_normal(n)
    X = O.rfft(x)
    Xn = np.fft.rfft(x)
    assert np.max(np.abs(X - Xn)) <= 1e-13 * np.max(np.abs(Xn))
    # c2r: unnormalised, Im of DC and Nyquist ignored (FFTW c2r)
    Y = Xn.copy()
    Y[0] += 3.0j
    Y[-1] += 5.0j
    xb = O.irfft(Y, n)
    assert np.max(np.abs(xb - n * x)) <= 1e-12 * n * np.max(np.abs(x))


def test_rfft2_roundtrip():
    x = RNG.standard_normal((6, 10))
    X = O.rfft2(x)
    assert np.max(np.abs(X - np.fft.rfft2(x))) <= 1e-13
    assert np.max(np.abs(O.irfft2(X, 10) - 60 * x)) <= 1e-12


def test_cspline_matches_scipy_natural():
    xs = np.sort(RNG.uniform(0, 10, 60))
    ys = np.sin(xs) * np.exp(xs / 4)
    xq = np.linspace(xs[0], xs[-1], 777)
    ref = si.CubicSpline(xs, ys, bc_type="natural")(xq)
    assert np.max(np.abs(O.cspline(xs, ys, xq) - ref)) <= 1e-12 * np.max(np.abs(ys))
    # natural spline reproduces straight lines exactly and returns NaN out of range
    assert np.allclose(O.cspline(xs, 2 * xs + 1, xq), 2 * xq + 1, rtol=0, atol=1e-12)
    assert np.isnan(O.cspline(xs, ys, [xs[0] - 1e-9])[0])


def test_akima_interior_and_linear():
    xs = np.sort(RNG.uniform(0, 10, 50))
    ys = np.cos(xs) * xs
    xq = np.linspace(xs[0], xs[-1], 1000)
    ref = si.Akima1DInterpolator(xs, ys)(xq)
    inner = (xq > xs[2]) & (xq < xs[-3])
    assert np.max(np.abs(O.akima(xs, ys, xq)[inner] - ref[inner])) <= 1e-12
    assert np.allclose(O.akima(xs, 3 * xs - 2, xq), 3 * xq - 2, rtol=0, atol=1e-12)
    # nodes are reproduced exactly
    assert np.array_equal(O.akima(xs, ys, xs[:-1]), ys[:-1])


def test_gauss_legendre_and_legendre_pl():
    for n in (5, 61, 256):
        x, w = O.gauss_legendre(n)
        nx, nw = np.polynomial.legendre.leggauss(n)
        assert np.max(np.abs(x - nx)) <= 2e-15 and np.max(np.abs(w - nw)) <= 1e-14
    for l in range(7):
        for v in (-0.93, -0.2, 0.0, 0.41, 1.0):
            ref = np.polynomial.legendre.legval(v, [0] * l + [1])
            assert abs(O.legendre_pl(l, v) - ref) <= 1e-15


def test_qag_closed_forms():
    r, st = O.qag_test(0, 0.0, 2.0, 1e-12)        # x^9 + cos x
    assert st == 0 and abs(r - (2 ** 10 / 10 + np.sin(2))) <= 1e-12 * r
    r, st = O.qag_test(1, 0.0, 10.0, 1e-10)       # sin(50x) e^{-x}, oscillatory
    exact = (50 - np.exp(-10) * (np.sin(500) + 50 * np.cos(500))) / 2501
    assert st == 0 and abs(r - exact) <= 1e-10 * abs(exact)
    r, st = O.qag_test(2, 0.0, 1.0, 1e-8)         # 1/sqrt(x): end-point singularity
    assert abs(r - 2.0) <= 1e-7


@pytest.mark.parametrize("l", range(6))
def test_spherical_bessel(l):
    mp.mp.dps = 30
    for x in (1e-8, 1e-3, 0.3, 0.49, 0.5, 0.7, 1.0, 2.0, 3.9, 4.5, 8.5, 31.4159, 270.0, 1e4):
        ref = float(mp.sqrt(mp.pi / (2 * x)) * mp.besselj(l + 0.5, x))
        assert abs(O.sph_jl(l, x) - ref) <= 1e-13 * abs(ref)


def test_threej_and_legendre_integral():
    from sympy.physics.wigner import wigner_3j
    for a in range(0, 5):
        for b in range(0, 5):
            for c in range(abs(a - b), a + b + 1):
                assert abs(O.threej000(a, b, c) - float(wigner_3j(a, b, c, 0, 0, 0))) <= 1e-14
    x, w = np.polynomial.legendre.leggauss(40)
    P = [np.polynomial.legendre.legval(x, [0] * l + [1]) for l in range(9)]
    for (n, m, a, b) in [(0, 0, 0, 0), (2, 2, 0, 0), (2, 4, 2, 4), (4, 4, 4, 4), (0, 2, 2, 0), (2, 2, 4, 2),
                         (1, 3, 2, 0)]:
        ref = np.sum(w * P[n] * P[m] * P[a] * P[b])
        assert abs(O.legendre_integral(n, m, a, b) - ref) <= 1e-14


def test_bicubic_reproduces_bilinear_and_separable_cubic():
    xs = np.sort(RNG.uniform(0, 3, 20))
    ys = np.sort(RNG.uniform(0, 2, 15))
    z = (1 + 2 * xs[None, :] - ys[:, None] + 0.5 * xs[None, :] * ys[:, None])   # z[j, i]
    from oracle.oracle import _f64  # noqa: F401
    import ctypes
    L = O.lib()
    L.orc_bicubic_new.restype = ctypes.c_void_p
    L.orc_bicubic_new.argtypes = [ctypes.c_size_t, ctypes.c_size_t, O.c_double_p, O.c_double_p, O.c_double_p]
    L.orc_bicubic_eval.restype = ctypes.c_double
    L.orc_bicubic_eval.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
    L.orc_bicubic_free.argtypes = [ctypes.c_void_p]
    zz = np.ascontiguousarray(z)
    h = L.orc_bicubic_new(len(xs), len(ys), xs.ctypes.data_as(O.c_double_p), ys.ctypes.data_as(O.c_double_p),
                          zz.ctypes.data_as(O.c_double_p))
    for _ in range(50):
        x, y = RNG.uniform(xs[0], xs[-1]), RNG.uniform(ys[0], ys[-1])
        assert abs(L.orc_bicubic_eval(h, x, y) - (1 + 2 * x - y + 0.5 * x * y)) <= 1e-12
    L.orc_bicubic_free(h)


def _bicubic(xs, ys, z):
    import ctypes
    L = O.lib()
    L.orc_bicubic_new.restype = ctypes.c_void_p
    L.orc_bicubic_new.argtypes = [ctypes.c_size_t, ctypes.c_size_t, O.c_double_p, O.c_double_p, O.c_double_p]
    L.orc_bicubic_eval.restype = ctypes.c_double
    L.orc_bicubic_eval.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
    L.orc_bicubic_free.argtypes = [ctypes.c_void_p]
    zz = np.ascontiguousarray(z)
    h = L.orc_bicubic_new(len(xs), len(ys), xs.ctypes.data_as(O.c_double_p), ys.ctypes.data_as(O.c_double_p),
                          zz.ctypes.data_as(O.c_double_p))
    return L, h, zz


def test_bicubic_separable_equals_product_of_natural_splines():
    """GSL's bicubic (common.c:579-613, 712-719) is a Hermite patch whose node derivatives
    are natural-cubic-spline derivatives along x, along y and (cross) along both.  For
    separable data g(x) h(y) those are S_g'(x_i) h_j, g_i S_h'(y_j) and S_g'(x_i) S_h'(y_j),
    and a cubic Hermite interpolant with spline node slopes is the spline itself, so the
    bicubic equals S_g(x) S_h(y) with S the natural cubic splines (scipy CubicSpline,
    bc_type='natural') -- end conditions, slope formula and cell lookup included."""
    from scipy.interpolate import CubicSpline
    rng = np.random.default_rng(77)
    xs = np.sort(rng.uniform(0.1, 3.0, 23))
    ys = np.sort(rng.uniform(-1.0, 2.0, 17))
    g = np.sin(2.3 * xs) + 0.3 * xs ** 2
    h = np.exp(-0.7 * ys) * np.cos(1.9 * ys)
    z = h[:, None] * g[None, :]          # z[j, i] = g(x_i) h(y_j)
    L, hnd, zz = _bicubic(xs, ys, z)
    Sg = CubicSpline(xs, g, bc_type="natural")
    Sh = CubicSpline(ys, h, bc_type="natural")
    pts = [(rng.uniform(xs[0], xs[-1]), rng.uniform(ys[0], ys[-1])) for _ in range(300)]
    pts += [(xs[0], ys[0]), (xs[-1], ys[-1]), (xs[3], ys[5]), (0.5 * (xs[0] + xs[1]), ys[-1])]
    for x, y in pts:
        ref = float(Sg(x) * Sh(y))
        assert abs(L.orc_bicubic_eval(hnd, x, y) - ref) <= 1e-12 * (1 + abs(ref)), (x, y)
    L.orc_bicubic_free(hnd)


def test_cspline_derivative_matches_scipy():
    """orc_cspline_deriv (the natural spline's first derivative, the node slopes of the
    bicubic) against scipy's CubicSpline(bc_type='natural').derivative(), at the nodes
    and between them."""
    from scipy.interpolate import CubicSpline
    L = O.lib()
    rng = np.random.default_rng(78)
    x = np.sort(rng.uniform(0, 5, 40))
    y = np.cos(1.7 * x) * np.exp(0.2 * x)
    c = np.empty(len(x))
    assert L.orc_cspline_init(len(x), x.ctypes.data_as(O.c_double_p), y.ctypes.data_as(O.c_double_p),
                              c.ctypes.data_as(O.c_double_p)) == 0
    dS = CubicSpline(x, y, bc_type="natural").derivative()
    for xq in list(x) + list(rng.uniform(x[0], x[-1], 200)):
        got = L.orc_cspline_deriv(len(x), x.ctypes.data_as(O.c_double_p), y.ctypes.data_as(O.c_double_p),
                                  c.ctypes.data_as(O.c_double_p), float(xq))
        assert abs(got - float(dS(xq))) <= 1e-12 * (1 + abs(float(dS(xq)))), xq


def test_window_2d_closed_form():
    """c_window_2d (utils.c:224-258) on an array of ones equals, entry by entry,
    w_row(i) w_col(j) with w(t) = t - sin(2 pi t)/(2 pi): columns j >= N2/2 - Ncut2 carry
    w((N2/2 - j)/Ncut2), rows N1/2 - a and N1/2 + 1 + a (0 <= a <= Ncut1) carry w(a/Ncut1),
    Ncut = floor(N/2 * width); every other entry is 1 (so the Nyquist column and rows
    N1/2, N1/2 + 1 vanish)."""
    w = lambda t: t - np.sin(2 * np.pi * t) / (2 * np.pi)  # noqa: E731
    for N1, N2, width in [(64, 48, 0.25), (40, 40, 0.25), (100, 64, 0.3)]:
        h1, h2 = N1 // 2, N2 // 2
        out = np.zeros((N1, h2 + 1, 2))
        out[..., 0] = 1.0
        out[..., 1] = -2.0
        O.lib().orc_window_2d(out.ctypes.data_as(O.c_double_p), width, h1, h2)
        c1, c2 = int(h1 * width), int(h2 * width)
        wr = np.ones(N1)
        for a in range(c1 + 1):
            wr[h1 - a] = w(a / c1)
            wr[h1 + 1 + a] = w(a / c1)
        wc = np.ones(h2 + 1)
        for j in range(h2 + 1):
            if h2 - j <= c2:
                wc[j] = w((h2 - j) / c2)
        exp = wr[:, None] * wc[None, :]
        assert np.allclose(out[..., 0], exp, rtol=0, atol=1e-15)
        assert np.allclose(out[..., 1], -2.0 * exp, rtol=0, atol=1e-15)
        assert np.all(out[:, h2, :] == 0) and np.all(out[h1, :, :] == 0) and np.all(out[h1 + 1, :, :] == 0)
