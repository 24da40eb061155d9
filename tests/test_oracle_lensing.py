"""The oracle's flat-sky lensing multipoles against the reference's fixtures
(tests/test_coffe.py:328-380 of the reference: lensing-lensing with flatsky_nonlocal,
density-lensing with flatsky_local_nonlocal and only the cross terms; rtol 5e-4 as
there), and the half-integer table against its closed form at one point."""
import os

import numpy as np
import pytest

import oracle as O
from oracle import lensing as OL

GOLD = os.path.join(os.path.dirname(__file__), "golden")
H = 0.67
H0h = H / 2997.92458
OMEGA_M = 0.3
N = 10000


@pytest.fixture(scope="module")
def inputs():
    from coffe_amd import background as B
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    bg = B.compute(omega_cdm=0.25, omega_baryon=0.05, h=H)
    return k * H, pk / H ** 3, bg


def _fixture(name, l):
    x, y = np.loadtxt(os.path.join(GOLD, f"benchmark_flatsky_{name}_multipoles{l}.dat")).T
    return x / H, y


@pytest.mark.parametrize("l", [0, 2, 4])
def test_lensing_lensing_vs_reference(inputs, l):
    k, pk, bg = inputs
    table = O.integral_table(1, l, k, pk, N)
    seps, ref = _fixture("lensing_lensing", l)
    got = [OL.flat_lensing_lensing(table, bg, 1.0, l, r, 0.0, 0.0, OMEGA_M, H0h) for r in seps]
    assert np.allclose(got, ref, rtol=5e-4, atol=0), (got, ref)


@pytest.mark.parametrize("l", [0, 2, 4])
def test_density_lensing_vs_reference(inputs, l):
    k, pk, bg = inputs
    tabs = [OL.half_table(l - 2 * kk + 3, l - 2 * kk + 1, k, pk, N) for kk in range(l // 2 + 1)]
    seps, ref = _fixture("density_lensing", l)
    got = [OL.flat_density_lensing(tabs, bg, 1.0, l, r, 0.0, 0.0, 1.0, 1.0, OMEGA_M, H0h) for r in seps]
    assert np.allclose(got, ref, rtol=5e-4, atol=0), (got, ref)


def test_odd_multipoles_vanish(inputs):
    k, pk, bg = inputs
    assert OL.flat_lensing_lensing(None, bg, 1.0, 3, 50.0, 0.0, 0.0, OMEGA_M, H0h) == 0.0
    assert OL.flat_density_lensing(None, bg, 1.0, 1, 50.0, 0.0, 0.0, 1.0, 1.0, OMEGA_M, H0h) == 0.0


def test_half_table_closed_form():
    """For P(k) = exp(-k^2 / 2) the table (3/2, 1/2) is r I(r) with
    I = int dk k^2 P j_{1/2}(kr) / (kr)^{3/2} / (2 pi^2); j_{1/2}(u) = sqrt(pi / (2u)) J_1(u)
    and int dk e^{-k^2/2} J_1(kr) = (1 - e^{-r^2/2}) / r give
    r I(r) = sqrt(pi/2) (1 - e^{-r^2/2}) / (2 pi^2 r^2)."""
    k = np.logspace(-5, 2.5, 4096)
    pk = np.exp(-k * k / 2)
    x, y = OL.half_table(3, 1, k, pk, 4096)
    sel = (x > 0.5) & (x < 10)
    want = np.sqrt(np.pi / 2) * (1 - np.exp(-x[sel] ** 2 / 2)) / (2 * np.pi ** 2 * x[sel] ** 2)
    assert np.allclose(y[sel], want, rtol=1e-5, atol=0), np.max(np.abs(y[sel] / want - 1))


def test_lensing_lensing_panel_convergence(inputs):
    """The oracle's brute-force x rule converges: 1024 and 2048 panels agree."""
    k, pk, bg = inputs
    table = O.integral_table(1, 2, k, pk, 4096)
    for r in (8.0, 160.0):
        a = OL.flat_lensing_lensing(table, bg, 0.7, 2, r, 0.1, 0.2, OMEGA_M, H0h, panels=1024)
        b = OL.flat_lensing_lensing(table, bg, 0.7, 2, r, 0.1, 0.2, OMEGA_M, H0h, panels=2048)
        assert abs(a - b) <= 1e-11 * abs(b), (r, a, b)


MUS = [0.0, 0.2, 0.5, 0.8, 0.95]


@pytest.fixture(scope="module")
def half_cf_table(inputs):
    k, pk, bg = inputs
    return OL.half_table(1, -1, k, pk, N)


@pytest.mark.parametrize("imu", range(5))
def test_lensing_lensing_corrfunc_vs_reference(inputs, half_cf_table, imu):
    """tests/test_corrfunc.c:360-420 of the reference (flatsky_nonlocal, mu = 0 ... 0.95)."""
    k, pk, bg = inputs
    x, y = np.loadtxt(os.path.join(GOLD, f"benchmark_flatsky_lensing_lensing_corrfunc{imu}.dat")).T
    got = [OL.flat_lensing_lensing_cf(half_cf_table, bg, 1.0, MUS[imu], r / H, 0.0, 0.0, OMEGA_M, H0h) for r in x]
    assert np.allclose(got, y, rtol=5e-4, atol=0), (got, y)


@pytest.mark.parametrize("l", [0, 2, 4])
def test_corrfunc_projects_onto_multipoles(inputs, half_cf_table, l):
    """(2l+1)/2 int dmu L_l(mu) xi(mu) of the correlation-function terms equals the
    multipole terms: lensing-lensing and density-lensing, each from tables of its own.
    Both terms are even in mu, so the projection is (2l+1) int_0^{pi/2} dtheta sin(theta)
    L_l(cos theta) xi(cos theta) with mu = cos theta, which removes the square-root
    behaviour of sqrt(1 - mu^2) at |mu| = 1 (96 Gauss-Legendre nodes, converged).  The
    sides read different FFTLog tables, each accurate to ~1e-4 against direct quadrature
    (the transform's window and aliasing, the same in the reference), so they agree to
    the reference tests' 5e-4 (measured: 2e-6 to 2.2e-4); a wrong factor, sign or table
    would be off by O(1)."""
    from numpy.polynomial.legendre import leggauss, legval
    k, pk, bg = inputs
    t, w = leggauss(96)
    th = (t + 1) * np.pi / 4
    wt = w * np.pi / 4 * np.sin(th)
    mus = np.cos(th)
    Ll = legval(mus, [0] * l + [1])
    table_1l = O.integral_table(1, l, k, pk, N)
    tabs = [OL.half_table(l - 2 * kk + 3, l - 2 * kk + 1, k, pk, N, small_r=False) for kk in range(l // 2 + 1)]
    s1, s2, b1, b2 = 0.2, -0.1, 1.3, 0.8
    for r in (30.0, 120.0):
        ll = [OL.flat_lensing_lensing_cf(half_cf_table, bg, 1.0, m, r, s1, s2, OMEGA_M, H0h, panels=256) for m in mus]
        dl = [OL.flat_density_lensing_cf(half_cf_table, bg, 1.0, m, r, s1, s2, b1, b2, OMEGA_M, H0h) for m in mus]
        proj_ll = (2 * l + 1) * np.sum(wt * Ll * ll)
        proj_dl = (2 * l + 1) * np.sum(wt * Ll * dl)
        want_ll = OL.flat_lensing_lensing(table_1l, bg, 1.0, l, r, s1, s2, OMEGA_M, H0h, panels=256)
        want_dl = OL.flat_density_lensing(tabs, bg, 1.0, l, r, s1, s2, b1, b2, OMEGA_M, H0h)
        assert abs(proj_ll - want_ll) <= 5e-4 * abs(want_ll), (l, r, proj_ll, want_ll)
        assert abs(proj_dl - want_dl) <= 5e-4 * abs(want_dl), (l, r, proj_dl, want_dl)
