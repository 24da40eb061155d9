"""ctypes bindings for oracle/liboracle.so (test infrastructure only).

Each wrapper names the reference function it restates; see the C sources for
file:line citations.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

c_double_p = ctypes.POINTER(ctypes.c_double)
c_size = ctypes.c_size_t


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = ctypes.CDLL(_LIB)
        _declare(_lib)
    return _lib


def _declare(L):
    d, p, s, i = ctypes.c_double, c_double_p, c_size, ctypes.c_int
    L.orc_fft.argtypes = [s, p, p, i]
    L.orc_rfft_r2c.argtypes = [s, p, p]
    L.orc_rfft_c2r.argtypes = [s, p, p]
    L.orc_rfft2_r2c.argtypes = [s, s, p, p]
    L.orc_rfft2_c2r.argtypes = [s, s, p, p]
    L.orc_cspline_interp.argtypes = [s, p, p, s, p, p]
    L.orc_akima_interp.argtypes = [s, p, p, s, p, p]
    L.orc_bicubic_interp.argtypes = [s, s, p, p, p, s, p, p, p]
    L.orc_gauss_legendre.argtypes = [s, p, p]
    L.orc_legendre_pl.argtypes = [i, d]
    L.orc_legendre_pl.restype = d
    L.orc_sph_jl.argtypes = [i, d]
    L.orc_sph_jl.restype = d
    L.orc_twofast_window.argtypes = [d, d, d, d, d, p]
    L.orc_twofast_lgamma.argtypes = [d, d, p]
    L.orc_twofast_lgamma.restype = None
    L.orc_twofast_mql.argtypes = [d, d, d, d, p]
    L.orc_twofast_selectqnu.argtypes = [d, d]
    L.orc_twofast_selectqnu.restype = d
    L.orc_twofast_1bessel.argtypes = [p, p, s, p, p, s, d, d, d, d, d, d]
    L.orc_integral_table.argtypes = [i, i, s, s, p, p, d, d, i, p, p]
    L.orc_divergent_tables.argtypes = [s, p, p, d, d, s, p, s, p, p, p, p]
    L.orc_multipoles.argtypes = [ctypes.POINTER(c_double_p), ctypes.POINTER(c_double_p), ctypes.POINTER(s), s,
                                 p, p, p, p, p, p, p, s, p, p, p, i, i, i, i, s, p, s,
                                 ctypes.POINTER(ctypes.c_int), s, p, i, d, p]
    L.orc_multipoles_dg.argtypes = [ctypes.POINTER(c_double_p), ctypes.POINTER(c_double_p), ctypes.POINTER(s), s,
                                    p, p, p, p, p, p, p, s, p, p, p, i, i, i, i, s, p, s,
                                    ctypes.POINTER(ctypes.c_int), s, p, i, d, ctypes.c_void_p, p]
    L.orc_corrfunc_dg.argtypes = [ctypes.POINTER(c_double_p), ctypes.POINTER(c_double_p), ctypes.POINTER(s), s,
                                  p, p, p, p, p, p, p, s, p, p, p, i, i, i, i, s, p, s, p, s, p, ctypes.c_void_p, p]
    L.orc_g_l.argtypes = [d, d, s, p, p]
    L.orc_g_l.restype = None
    L.orc_window_2d.argtypes = [p, d, ctypes.c_long, ctypes.c_long]
    L.orc_window_2d.restype = None
    L.orc_cspline_init.argtypes = [s, p, p, p]
    L.orc_cspline_deriv.argtypes = [s, p, p, p, d]
    L.orc_cspline_deriv.restype = d
    L.orc_two_sph_bessel.argtypes = [p, p, p, ctypes.c_long, ctypes.c_long, d, d, d, d, d, p, p, p]
    L.orc_covariance_fftlog_integral.argtypes = [s, p, p, d, i, i, s, s, p, d, d, p]
    L.orc_threej000.argtypes = [i, i, i]
    L.orc_threej000.restype = d
    L.orc_legendre_integral.argtypes = [i, i, i, i]
    L.orc_legendre_integral.restype = d
    L.orc_covariance_coefficient.argtypes = [d, d, d, i]
    L.orc_covariance_coefficient.restype = d
    L.orc_covariance_assemble.argtypes = [s, s, ctypes.POINTER(ctypes.c_int), s, p, ctypes.POINTER(ctypes.c_int),
                                          p, p, p, p, p, p, p, i, i, i, i, p, p, p]
    ll = ctypes.POINTER(ctypes.c_longlong)
    L.orc_set_threads.argtypes = [i]
    L.orc_reference_batch.argtypes = [i, ll, p, p, ll, s, p, p, p, p, p, p, p, s, p, p, p, i, i, i, i, s, p, s,
                                      ctypes.POINTER(ctypes.c_int), s, p, i, d, p]
    L.orc_tables_at_sep_batch.argtypes = [i, ll, p, p, ll, i, ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int), s, p, p]
    L.orc_tables_at_sep_jvp.argtypes = [i, ll, p, p, p, ll, i, ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int), s, p, p, p]
    L.orc_akima_jvp.argtypes = [s, p, p, p, s, p, p, p]
    L.orc_mu_weights_hp.argtypes = [s, p, p, p, p, p, p, p, s, p, p, p, i, i, i, i, s, p, s,
                                    ctypes.POINTER(ctypes.c_int), s, p, s, p, p, p]
    L.orc_mu_weights_hp_dg.argtypes = [s, p, p, p, p, p, p, p, s, p, p, p, i, i, i, i, s, p, s,
                                       ctypes.POINTER(ctypes.c_int), s, p, s, p, p, ctypes.c_void_p, p, p]
    L.orc_qag_test.argtypes = [i, d, d, d, ctypes.POINTER(ctypes.c_int)]
    L.orc_qag_test.restype = d


def _p(a):
    return a.ctypes.data_as(c_double_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class OracleError(RuntimeError):
    pass


def _check(st, what):
    if st != 0:
        raise OracleError(f"{what} failed with oracle status {st}")


def fft(x, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    _check(lib().orc_fft(len(x), x.ctypes.data_as(c_double_p), out.ctypes.data_as(c_double_p), sign), "fft")
    return out


def rfft(x):
    """FFTW r2c (unnormalised)."""
    x = _f64(x)
    out = np.empty(len(x) // 2 + 1, dtype=np.complex128)
    _check(lib().orc_rfft_r2c(len(x), _p(x), out.ctypes.data_as(c_double_p)), "rfft")
    return out


def irfft(X, n):
    """FFTW c2r (unnormalised; Im of DC and Nyquist ignored)."""
    X = np.ascontiguousarray(X, dtype=np.complex128)
    out = np.empty(n)
    _check(lib().orc_rfft_c2r(n, X.ctypes.data_as(c_double_p), _p(out)), "irfft")
    return out


def rfft2(x):
    x = _f64(x)
    n1, n2 = x.shape
    out = np.empty((n1, n2 // 2 + 1), dtype=np.complex128)
    _check(lib().orc_rfft2_r2c(n1, n2, _p(x), out.ctypes.data_as(c_double_p)), "rfft2")
    return out


def irfft2(X, n2):
    X = np.ascontiguousarray(X, dtype=np.complex128)
    n1 = X.shape[0]
    out = np.empty((n1, n2))
    _check(lib().orc_rfft2_c2r(n1, n2, X.ctypes.data_as(c_double_p), _p(out)), "irfft2")
    return out


def cspline(x, y, xq):
    x, y, xq = _f64(x), _f64(y), _f64(np.atleast_1d(xq))
    out = np.empty(len(xq))
    _check(lib().orc_cspline_interp(len(x), _p(x), _p(y), len(xq), _p(xq), _p(out)), "cspline")
    return out


def akima(x, y, xq):
    x, y, xq = _f64(x), _f64(y), _f64(np.atleast_1d(xq))
    out = np.empty(len(xq))
    _check(lib().orc_akima_interp(len(x), _p(x), _p(y), len(xq), _p(xq), _p(out)), "akima")
    return out


def bicubic(x, y, z, xq, yq):
    """GSL's bicubic interpolant of z[j, i] = f(x_i, y_j) (natural-spline node
    derivatives) at the points (xq, yq); NaN outside the grid."""
    x, y, z = _f64(x), _f64(y), _f64(z)
    xq, yq = np.broadcast_arrays(_f64(np.atleast_1d(xq)), _f64(np.atleast_1d(yq)))
    xq, yq = _f64(xq.ravel()), _f64(yq.ravel())
    out = np.empty(len(xq))
    _check(lib().orc_bicubic_interp(len(x), len(y), _p(x), _p(y), _p(z), len(xq), _p(xq), _p(yq), _p(out)),
           "bicubic")
    return out


def gauss_legendre(n):
    x, w = np.empty(n), np.empty(n)
    lib().orc_gauss_legendre(n, _p(x), _p(w))
    return x, w


def legendre_pl(l, x):
    return lib().orc_legendre_pl(l, x)


def sph_jl(l, x):
    return lib().orc_sph_jl(l, x)


def qag_test(kind, a, b, epsrel):
    st = ctypes.c_int(0)
    r = lib().orc_qag_test(kind, a, b, epsrel, ctypes.byref(st))
    return r, st.value


def twofast_window(value, xmin, xmax, xleft, xright):
    out = ctypes.c_double()
    _check(lib().orc_twofast_window(value, xmin, xmax, xleft, xright, ctypes.byref(out)), "twofast_window")
    return out.value


def twofast_lgamma(z):
    out = (ctypes.c_double * 2)()
    lib().orc_twofast_lgamma(z.real, z.imag, out)
    return complex(out[0], out[1])


def twofast_mql(t, q, l, alpha):
    out = (ctypes.c_double * 2)()
    _check(lib().orc_twofast_mql(t, q, l, alpha, out), "twofast_mql")
    return complex(out[0], out[1])


def twofast_selectqnu(l, nu):
    return lib().orc_twofast_selectqnu(l, nu)


def twofast_1bessel(k, pk, N, l, nu, r0=1.0, k0=None, kmin=None, kmax=None):
    """twofast.c:362-483; returns (r, y)."""
    k, pk = _f64(k), _f64(pk)
    kmin = k[0] if kmin is None else kmin
    kmax = k[-1] if kmax is None else kmax
    k0 = kmin if k0 is None else k0
    ox, oy = np.empty(N), np.empty(N)
    _check(lib().orc_twofast_1bessel(_p(ox), _p(oy), N, _p(k), _p(pk), len(k), float(l), float(nu), r0, k0, kmin,
                                     kmax), "twofast_1bessel")
    return ox, oy


def integral_table(n, l, k, pk, N, small_r=True):
    """integrals.c:459-573; returns the (N + 43)-point table (x, y)."""
    k, pk = _f64(k), _f64(pk)
    ox, oy = np.empty(N + 43), np.empty(N + 43)
    _check(lib().orc_integral_table(n, l, N, len(k), _p(k), _p(pk), k[0], k[-1], int(small_r), _p(ox), _p(oy)),
           "integral_table")
    return ox, oy


def divergent_tables(k, pk, sep, chi, kmin=None, kmax=None):
    """integrals.c:772-909: (I_0^4(sep) sep^4, the r = 0 renormalisation at sep (times
    sep^4), the chi x chi renormalisation grid [a, b] = (chi_b, chi_a)), all / 2 pi^2."""
    k, pk, sep, chi = _f64(k), _f64(pk), _f64(sep), _f64(chi)
    kmin = k[0] if kmin is None else kmin
    kmax = k[-1] if kmax is None else kmax
    i40, r0 = np.empty(len(sep)), np.empty(len(sep))
    ren = np.empty((len(chi), len(chi)))
    _check(lib().orc_divergent_tables(len(k), _p(k), _p(pk), kmin, kmax, len(sep), _p(sep), len(chi), _p(chi),
                                      _p(i40), _p(r0), _p(ren)), "divergent_tables")
    return i40, r0, ren


TABLE_ORDER = [(0, 0), (0, 2), (0, 4), (1, 1), (1, 3), (2, 0), (2, 2), (3, 1)]  # (n, l), integrals.c:599-608


class DG(ctypes.Structure):
    """orc_dg: the d2 / g1-g3 inputs (functions.c:574-669, 856-903, 1144-1428)."""
    _fields_ = [("d2", ctypes.c_int), ("g1", ctypes.c_int), ("g2", ctypes.c_int), ("g3", ctypes.c_int),
                ("H0h", ctypes.c_double), ("Omega_m", ctypes.c_double), ("bg_a", c_double_p),
                ("nmb", c_size), ("mb_z", c_double_p), ("s1", c_double_p), ("s2", c_double_p),
                ("fevo1", c_double_p), ("fevo2", c_double_p), ("ndiv", c_size), ("div_x", c_double_p),
                ("div_y", c_double_p), ("nren", c_size), ("ren_chi", c_double_p), ("ren", c_double_p),
                ("only_cross", ctypes.c_int)]


def make_dg(flags, H0h, Omega_m, bg_a, mb_z, s1, s2, fevo1, fevo2, div_x, div_y, ren_chi, ren, only_cross=0):
    """An orc_dg for the flags dict (d2, g1, g2, g3); only_cross leaves their autos out
    (only_cross_correlations); returns (struct, the arrays it points into, which the
    caller keeps alive)."""
    keep = [_f64(a) for a in (bg_a, mb_z, s1, s2, fevo1, fevo2, div_x, div_y, ren_chi, ren)]
    st = DG(int(flags.get("d2", 0)), int(flags.get("g1", 0)), int(flags.get("g2", 0)), int(flags.get("g3", 0)),
            float(H0h), float(Omega_m), _p(keep[0]), len(keep[1]), _p(keep[1]), _p(keep[2]), _p(keep[3]),
            _p(keep[4]), _p(keep[5]), len(keep[6]), _p(keep[6]), _p(keep[7]), len(keep[8]), _p(keep[8]),
            _p(keep[9]), int(only_cross))
    return st, keep


def multipoles(tables, bg, bias_z, bias1, bias2, z_mean, ls, sep, den=True, rsd=True, d1=False, flatsky=False,
               quad=0, epsrel=1e-5, dg=None):
    """functions.c:89-1484 + signal.c:83-111,521-542 + multipoles.c:48-136.

    tables: dict (n, l) -> (x, y); bg: dict of arrays on bg['z'] (chi, D1, f, cH, G1, G2).
    dg: an orc_dg (make_dg) for the d2 / g1-g3 terms, or None.
    Returns array (len(z_mean), len(ls), len(sep)).
    """
    T = len(TABLE_ORDER)
    keep = []
    xs = (c_double_p * T)()
    ys = (c_double_p * T)()
    ns = (c_size * T)()
    for t, key in enumerate(TABLE_ORDER):
        if key in tables:
            x, y = _f64(tables[key][0]), _f64(tables[key][1])
            keep += [x, y]
            xs[t], ys[t], ns[t] = _p(x), _p(y), len(x)
        else:
            xs[t], ys[t], ns[t] = None, None, 0
    arrs = {kk: _f64(bg[kk]) for kk in ("z", "chi", "D1", "f", "cH", "G1", "G2")}
    bz, b1, b2 = _f64(bias_z), _f64(bias1), _f64(bias2)
    zm, sp = _f64(np.atleast_1d(z_mean)), _f64(np.atleast_1d(sep))
    lsa = (ctypes.c_int * len(ls))(*ls)
    out = np.empty(len(zm) * len(ls) * len(sp))
    _check(lib().orc_multipoles_dg(xs, ys, ns, len(arrs["z"]), _p(arrs["z"]), _p(arrs["chi"]), _p(arrs["D1"]),
                                   _p(arrs["f"]), _p(arrs["cH"]), _p(arrs["G1"]), _p(arrs["G2"]), len(bz), _p(bz),
                                   _p(b1), _p(b2), int(den), int(rsd), int(d1), int(flatsky), len(zm), _p(zm),
                                   len(ls), lsa, len(sp), _p(sp), int(quad), float(epsrel),
                                   ctypes.byref(dg) if dg is not None else None, _p(out)), "multipoles")
    return out.reshape(len(zm), len(ls), len(sp))


def corrfunc(tables, bg, bias_z, bias1, bias2, z_mean, mu, sep, den=True, rsd=True, d1=False, flatsky=False,
             dg=None):
    """The local correlation function (functions.c:89-1484) at every (z_mean, mu, sep), the
    inputs of multipoles(); returns array (len(z_mean), len(mu), len(sep))."""
    T = len(TABLE_ORDER)
    keep = []
    xs = (c_double_p * T)()
    ys = (c_double_p * T)()
    ns = (c_size * T)()
    for t, key in enumerate(TABLE_ORDER):
        if key in tables:
            x, y = _f64(tables[key][0]), _f64(tables[key][1])
            keep += [x, y]
            xs[t], ys[t], ns[t] = _p(x), _p(y), len(x)
        else:
            xs[t], ys[t], ns[t] = None, None, 0
    arrs = {kk: _f64(bg[kk]) for kk in ("z", "chi", "D1", "f", "cH", "G1", "G2")}
    bz, b1, b2 = _f64(bias_z), _f64(bias1), _f64(bias2)
    zm, mu_a, sp = _f64(np.atleast_1d(z_mean)), _f64(np.atleast_1d(mu)), _f64(np.atleast_1d(sep))
    out = np.empty(len(zm) * len(mu_a) * len(sp))
    _check(lib().orc_corrfunc_dg(xs, ys, ns, len(arrs["z"]), _p(arrs["z"]), _p(arrs["chi"]), _p(arrs["D1"]),
                                 _p(arrs["f"]), _p(arrs["cH"]), _p(arrs["G1"]), _p(arrs["G2"]), len(bz), _p(bz),
                                 _p(b1), _p(b2), int(den), int(rsd), int(d1), int(flatsky), len(zm), _p(zm),
                                 len(mu_a), _p(mu_a), len(sp), _p(sp), ctypes.byref(dg) if dg is not None else None,
                                 _p(out)), "corrfunc")
    return out.reshape(len(zm), len(mu_a), len(sp))


def g_l(l, nu, eta):
    eta = _f64(eta)
    out = np.empty(len(eta), dtype=np.complex128)
    lib().orc_g_l(l, nu, len(eta), _p(eta), out.ctypes.data_as(c_double_p))
    return out


def two_sph_bessel(k1, k2, fk, l1, l2, nu1=1.01, nu2=1.01, width=0.25):
    k1, k2, fk = _f64(k1), _f64(k2), _f64(fk)
    N1, N2 = len(k1), len(k2)
    r1, r2, res = np.empty(N1), np.empty(N2), np.empty((N1, N2))
    _check(lib().orc_two_sph_bessel(_p(k1), _p(k2), _p(fk), N1, N2, l1, l2, nu1, nu2, width, _p(r1), _p(r2),
                                    _p(res)), "two_sph_bessel")
    return r1, r2, res


def covariance_fftlog_integral(k, pk, power, l1, l2, N, sep, kmin=None, kmax=None):
    """covariance.c:357-470; returns R x R array indexed [n, m] (value at sep_m, sep_n)."""
    k, pk, sep = _f64(k), _f64(pk), _f64(sep)
    kmin = k[0] if kmin is None else kmin
    kmax = k[-1] if kmax is None else kmax
    R = len(sep)
    out = np.empty(R * R)
    _check(lib().orc_covariance_fftlog_integral(len(k), _p(k), _p(pk), power, l1, l2, N, R, _p(sep), kmin, kmax,
                                                _p(out)), "covariance_fftlog_integral")
    return out.reshape(R, R)


def threej000(a, b, c):
    return lib().orc_threej000(a, b, c)


def legendre_integral(n, m, a, b):
    return lib().orc_legendre_integral(n, m, a, b)


def covariance_assemble(ls, sep, bias, f, D1, nbar, vol, pix, pk_int, pk2_int, poisson=True, mixed=True,
                        cosmic=True, pops=None, nbar2=None, window=False):
    """covariance.c:790-1060.  bias: (nz,) (one value for all four biases) or (nz, 4)
    (galaxy biases 1..4 per bin); pops: covariance_pop1..4 (default 1, 1, 1, 1);
    nbar / nbar2: density1 / density2 per bin (nbar2 defaults to nbar)."""
    bias = np.asarray(bias, dtype=np.float64)
    if bias.ndim == 1:
        bias = np.repeat(bias[:, None], 4, axis=1)
    nz, nl, R = len(bias), len(ls), len(sep)
    lsa = (ctypes.c_int * nl)(*ls)
    pa = (ctypes.c_int * 4)(*(pops if pops is not None else (1, 1, 1, 1)))
    nbar2 = nbar if nbar2 is None else nbar2
    arrs = [_f64(np.atleast_1d(a)) for a in (sep, bias, f, D1, nbar, nbar2, vol, pix)]
    pk_int, pk2_int = _f64(pk_int), _f64(pk2_int)
    out = np.empty(nz * nl * R * nl * R)
    p = [_p(a) for a in arrs]
    _check(lib().orc_covariance_assemble(nz, nl, lsa, R, p[0], pa, *p[1:], int(poisson), int(mixed),
                                         int(cosmic), int(window), _p(pk_int), _p(pk2_int), _p(out)),
           "covariance_assemble")
    return out.reshape(nz, nl, R, nl, R)


def set_threads(n):
    """OpenMP thread count of the oracle's parallel loops; returns the previous one."""
    return int(lib().orc_set_threads(int(n)))


def reference_batch(specs, nfft, bg, bias_z, bias1, bias2, z_mean, ls, sep, den=True, rsd=True, d1=False,
                    flatsky=True, small_r=True, epsrel=1e-5):
    """CPU baseline: reference sequence per spectrum (8 tables + QAG multipoles), OpenMP over spectra.
    specs: list of (k, P); returns (nspec, nz, nl, nsep)."""
    off = np.concatenate([[0], np.cumsum([len(s[0]) for s in specs])]).astype(np.longlong)
    k = _f64(np.concatenate([s[0] for s in specs]))
    pk = _f64(np.concatenate([s[1] for s in specs]))
    nf = np.ascontiguousarray(nfft, dtype=np.longlong)
    arrs = {kk: _f64(bg[kk]) for kk in ("z", "chi", "D1", "f", "cH", "G1", "G2")}
    bz, b1, b2 = _f64(bias_z), _f64(bias1), _f64(bias2)
    zm, sp = _f64(np.atleast_1d(z_mean)), _f64(np.atleast_1d(sep))
    lsa = (ctypes.c_int * len(ls))(*ls)
    out = np.empty(len(specs) * len(zm) * len(ls) * len(sp))
    llp = ctypes.POINTER(ctypes.c_longlong)
    _check(lib().orc_reference_batch(len(specs), off.ctypes.data_as(llp), _p(k), _p(pk), nf.ctypes.data_as(llp),
                                     len(arrs["z"]), _p(arrs["z"]), _p(arrs["chi"]), _p(arrs["D1"]), _p(arrs["f"]),
                                     _p(arrs["cH"]), _p(arrs["G1"]), _p(arrs["G2"]), len(bz), _p(bz), _p(b1), _p(b2),
                                     int(den), int(rsd), int(d1), int(flatsky), len(zm), _p(zm), len(ls), lsa,
                                     len(sp), _p(sp), int(small_r), float(epsrel), _p(out)), "reference_batch")
    return out.reshape(len(specs), len(zm), len(ls), len(sp))


def tables_at_sep_batch(specs, nfft, tables, sep):
    """Akima of each oracle table at sep, for a list of (k, P); returns (nspec, ntab, nsep)."""
    off = np.concatenate([[0], np.cumsum([len(s[0]) for s in specs])]).astype(np.longlong)
    k = _f64(np.concatenate([s[0] for s in specs]))
    pk = _f64(np.concatenate([s[1] for s in specs]))
    nf = np.ascontiguousarray(nfft, dtype=np.longlong)
    tn = (ctypes.c_int * len(tables))(*[t[0] for t in tables])
    tl = (ctypes.c_int * len(tables))(*[t[1] for t in tables])
    sp = _f64(sep)
    out = np.empty(len(specs) * len(tables) * len(sp))
    llp = ctypes.POINTER(ctypes.c_longlong)
    _check(lib().orc_tables_at_sep_batch(len(specs), off.ctypes.data_as(llp), _p(k), _p(pk), nf.ctypes.data_as(llp),
                                         len(tables), tn, tl, len(sp), _p(sp), _p(out)), "tables_at_sep_batch")
    return out.reshape(len(specs), len(tables), len(sp))


def akima_jvp(x, y, dy, xq):
    """Akima interpolant at xq and its derivative along dy of the node values."""
    x, y, dy, xq = _f64(x), _f64(y), _f64(dy), _f64(np.atleast_1d(xq))
    out, dout = np.empty(len(xq)), np.empty(len(xq))
    _check(lib().orc_akima_jvp(len(x), _p(x), _p(y), _p(dy), len(xq), _p(xq), _p(out), _p(dout)), "akima_jvp")
    return out, dout


def tables_at_sep_jvp(specs, nfft, tables, sep, dirs):
    """orc_tables_at_sep_batch and its forward-mode derivative along dirs (one array per
    spectrum, like its P); returns (values, tangents), each (nspec, ntab, nsep)."""
    off = np.concatenate([[0], np.cumsum([len(s[0]) for s in specs])]).astype(np.longlong)
    k = _f64(np.concatenate([s[0] for s in specs]))
    pk = _f64(np.concatenate([s[1] for s in specs]))
    v = _f64(np.concatenate(dirs))
    nf = np.ascontiguousarray(nfft, dtype=np.longlong)
    tn = (ctypes.c_int * len(tables))(*[t[0] for t in tables])
    tl = (ctypes.c_int * len(tables))(*[t[1] for t in tables])
    sp = _f64(sep)
    out = np.empty(len(specs) * len(tables) * len(sp))
    dout = np.empty_like(out)
    llp = ctypes.POINTER(ctypes.c_longlong)
    _check(lib().orc_tables_at_sep_jvp(len(specs), off.ctypes.data_as(llp), _p(k), _p(pk), _p(v),
                                       nf.ctypes.data_as(llp), len(tables), tn, tl, len(sp), _p(sp), _p(out),
                                       _p(dout)), "tables_at_sep_jvp")
    shape = (len(specs), len(tables), len(sp))
    return out.reshape(shape), dout.reshape(shape)


def mu_weights_hp(bg, bias_z, bias1, bias2, z_mean, ls, sep, den=True, rsd=True, d1=False, flatsky=False, nmu=256,
                  nodes=None, dg=None):
    """The reference's local-term weights A[z][l][t][r] (t: the 8 standard tables, and I_0^4
    as a ninth with dg) evaluated in binary128 under an nmu-point Gauss-Legendre rule (or
    the given (mu, w) nodes); see orc_mu_weights_hp_dg.  With dg (make_dg) returns (A, ren),
    ren [z][l][r] the renormalisation grid's part of the multipoles."""
    arrs = {kk: _f64(bg[kk]) for kk in ("z", "chi", "D1", "f", "cH", "G1", "G2")}
    bz, b1, b2 = _f64(bias_z), _f64(bias1), _f64(bias2)
    zm, sp = _f64(np.atleast_1d(z_mean)), _f64(np.atleast_1d(sep))
    if nodes is None:
        mu, w = np.polynomial.legendre.leggauss(nmu)
    else:
        mu, w = nodes
    mu, w = _f64(mu), _f64(w)
    lsa = (ctypes.c_int * len(ls))(*ls)
    nt = 9 if dg is not None else 8
    out = np.empty(len(zm) * len(ls) * nt * len(sp))
    ren = np.zeros(len(zm) * len(ls) * len(sp))
    _check(lib().orc_mu_weights_hp_dg(len(arrs["z"]), _p(arrs["z"]), _p(arrs["chi"]), _p(arrs["D1"]), _p(arrs["f"]),
                                      _p(arrs["cH"]), _p(arrs["G1"]), _p(arrs["G2"]), len(bz), _p(bz), _p(b1),
                                      _p(b2), int(den), int(rsd), int(d1), int(flatsky), len(zm), _p(zm), len(ls),
                                      lsa, len(sp), _p(sp), len(mu), _p(mu), _p(w),
                                      ctypes.byref(dg) if dg is not None else None, _p(out), _p(ren)),
           "mu_weights_hp")
    A = out.reshape(len(zm), len(ls), nt, len(sp))
    if dg is None:
        return A
    return A, ren.reshape(len(zm), len(ls), len(sp))
