"""The oracle's divergent table I_0^4 and its renormalisations (integrals.c:772-909),
pinned by the reference's own fixtures and by closed forms.

The fixtures (test_integrals.c:155-245) were made with the default parameters: the
embedded CLASS spectrum in its own units (k in h/Mpc, P in (Mpc/h)^3 as in
PkL_CLASS.dat), k_min = 1e-5, k_max = 300 (parser.c:840-841), x = sep H0 and
values times H0^4.
"""
import os

import numpy as np
from scipy.special import erf

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
COFFE_H0 = 1 / 2997.92458
KMIN, KMAX = 1e-5, 300.0


def _class_spectrum():
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    return k, pk


def _ren_fixture():
    x1, x2, v = np.loadtxt(os.path.join(GOLD, "benchmark_integral8_renormalization_every10.dat")).T
    n = int(round(np.sqrt(len(v))))
    return x1.reshape(n, n)[:, 0] / COFFE_H0, v.reshape(n, n) / COFFE_H0 ** 4


def test_i40_matches_reference_fixture():
    """All 297 rows of benchmark_integral8.dat: the reference's separation list and
    I_0^4(sep) sep^4 (the fixture has 11 digits; the oracle's QAG agrees to ~1e-10)."""
    k, pk = _class_spectrum()
    x, y = np.loadtxt(os.path.join(GOLD, "benchmark_integral8.dat")).T
    sep = x / COFFE_H0
    i40, r0, _ = O.divergent_tables(k, pk, sep, np.zeros(0), kmin=KMIN, kmax=KMAX)
    assert np.max(np.abs(i40 * COFFE_H0 ** 4 / y - 1)) < 1e-8
    assert np.all(np.isfinite(r0))


def test_renormalization_matches_reference_fixture():
    """A 5 x 5 subgrid of the 201 x 201 renormalisation grid (7-digit fixture)."""
    k, pk = _class_spectrum()
    chi, ref = _ren_fixture()
    pick = np.arange(0, len(chi), 5)
    _, _, ren = O.divergent_tables(k, pk, np.zeros(0), chi[pick], kmin=KMIN, kmax=KMAX)
    assert np.max(np.abs(ren / ref[np.ix_(pick, pick)] - 1)) < 1e-5
    assert np.array_equal(ren, ren.T)


def _gauss_F(c, R):
    # int_0^inf e^{-R^2 x^2} (1 - cos c x) / x^2 dx
    return np.pi / 2 * (c * erf(c / (2 * R)) + 2 * R / np.sqrt(np.pi) * (np.exp(-c * c / (4 * R * R)) - 1))


def _gauss_jj(a, b, R):
    # int_0^inf e^{-R^2 k^2} j_0(k a) j_0(k b) dk
    if a == 0 and b == 0:
        return np.sqrt(np.pi) / (2 * R)
    if a == 0 or b == 0:
        c = max(a, b)
        return np.pi / (2 * c) * erf(c / (2 * R))
    return (_gauss_F(a + b, R) - _gauss_F(abs(a - b), R)) / (2 * a * b)


def test_gaussian_closed_forms():
    """P = k^2 exp(-k^2 R^2): I_0^4 sep^4 = (pi / 2 sep) erf(sep / 2R), the grid is the
    erf / Gaussian closed form of int e^{-k^2 R^2} j_0 j_0 dk, and the r = 0
    renormalisation is sep^4 (sqrt(pi)/2R - grid(sep, sep)); all / 2 pi^2."""
    R = 8.0
    k = np.logspace(-8, np.log10(4.0), 4000)
    pk = k * k * np.exp(-(k * R) ** 2)
    sep = np.array([1.0, 5.0, 20.0, 80.0, 300.0])
    chi = np.array([0.0, 3.0, 17.0, 60.0, 250.0])
    i40, r0, ren = O.divergent_tables(k, pk, sep, chi)
    tp = 2 * np.pi ** 2
    want = np.pi / (2 * sep) * erf(sep / (2 * R)) / tp
    assert np.max(np.abs(i40 / want - 1)) < 2e-5
    w0 = sep ** 4 * (np.sqrt(np.pi) / (2 * R) - np.array([_gauss_jj(s, s, R) for s in sep])) / tp
    assert np.max(np.abs(r0 / w0 - 1)) < 2e-5
    wg = np.array([[_gauss_jj(b, a, R) for b in chi] for a in chi]) / tp
    assert np.max(np.abs(ren / wg - 1)) < 2e-5


def test_separation_list_is_the_fixtures():
    """The engine's default separation list is the fixture's x / H0 (integrals.c:76-99)."""
    from coffe_amd.engine import COFFE_SEP
    x, _ = np.loadtxt(os.path.join(GOLD, "benchmark_integral8.dat")).T
    assert len(COFFE_SEP) == len(x) == 297
    assert np.max(np.abs(COFFE_SEP / (x / COFFE_H0) - 1)) < 1e-9
