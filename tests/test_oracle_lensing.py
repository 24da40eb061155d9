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
