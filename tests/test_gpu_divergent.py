"""GPU divergent table I_0^4 and its renormalisations (k_divergent, quad_kernels.hip)
against the reference's fixtures, the oracle and closed forms (integrals.c:772-909)."""
import os

import numpy as np
import pytest
from scipy.special import erf

import oracle as O
from tests.synth import spectrum
from tests.test_oracle_divergent import COFFE_H0, KMAX, KMIN, _class_spectrum, _gauss_jj, _ren_fixture

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_fixtures_full_grid(engine):
    """The reference's configuration: the 297 separations and the 201 x 201 grid on
    [0, chi_max]; I_0^4 against benchmark_integral8.dat (11 digits), the grid against
    every 10th node of benchmark_integral8_renormalization.dat (7 digits); the grid is
    symmetric bit for bit and no integral reaches the interval cap."""
    from coffe_amd.engine import COFFE_SEP, divergent_tables
    k, pk = _class_spectrum()
    chi10, ref = _ren_fixture()
    chi = np.linspace(0.0, chi10[-1], 201)
    print("divergent tables: 297 seps, 201 x 201 grid, chi_max", chi[-1], flush=True)
    sep, i40, r0, ren, status, nint = divergent_tables(k, pk, chi=chi, kmin=KMIN, kmax=KMAX, return_status=True)
    ns = len(sep)
    # the reference discards the r = 0 renormalisation at its smallest separation
    # (integrals.c:838-839); it is a cancellation 1 - j0^2 ~ 1e-13 there
    assert not np.any(status[:ns] == 1) and not np.any(status[2 * ns:] == 1)
    assert np.all(nint >= 1)
    x, y = np.loadtxt(os.path.join(GOLD, "benchmark_integral8.dat")).T
    assert np.array_equal(sep, COFFE_SEP)
    assert np.max(np.abs(i40 * COFFE_H0 ** 4 / y - 1)) < 1e-8
    assert np.array_equal(ren, ren.T)
    assert np.max(np.abs(ren[::10, ::10] / ref - 1)) < 1e-5
    # r0(s) / s^4 + grid(s, s) = grid(0, 0) = int P / k^2 (1 - j0^2 + j0^2 = 1), to the
    # QAG tolerance, on a grid through some of the separations
    s5 = np.array([0.0, 30.0, 100.0, 500.0, 2000.0, 6000.0])
    _, _, r5, g5 = divergent_tables(k, pk, sep=s5[1:], chi=s5, kmin=KMIN, kmax=KMAX)
    assert np.max(np.abs((r5 / s5[1:] ** 4 + np.diag(g5)[1:]) / g5[0, 0] - 1)) < 5e-5


def test_against_oracle(engine):
    """A synthetic spectrum: every separation and a 9 x 9 grid against the oracle (the
    same QAG bisection sequence)."""
    from coffe_amd.engine import divergent_tables
    k, pk = spectrum(4096)
    chi = np.linspace(0.0, 6000.0, 9)
    print("divergent tables vs oracle", flush=True)
    sep, i40, r0, ren, status, _ = divergent_tables(k, pk, chi=chi, return_status=True)
    oi, orr, oren = O.divergent_tables(k, pk, sep, chi)
    assert np.max(np.abs(i40 / oi - 1)) < 1e-12
    assert np.max(np.abs(ren / oren - 1)) < 1e-12
    assert np.max(np.abs(r0[1:] / orr[1:] - 1)) < 1e-8


def test_gaussian_closed_forms(engine):
    from coffe_amd.engine import divergent_tables
    R = 8.0
    k = np.logspace(-8, np.log10(4.0), 4000)
    pk = k * k * np.exp(-(k * R) ** 2)
    sep = np.array([1.0, 5.0, 20.0, 80.0, 300.0])
    chi = np.array([0.0, 3.0, 17.0, 60.0, 250.0])
    print("divergent tables: Gaussian closed forms", flush=True)
    _, i40, r0, ren = divergent_tables(k, pk, sep=sep, chi=chi)
    tp = 2 * np.pi ** 2
    assert np.max(np.abs(i40 / (np.pi / (2 * sep) * erf(sep / (2 * R)) / tp) - 1)) < 2e-5
    w0 = sep ** 4 * (np.sqrt(np.pi) / (2 * R) - np.array([_gauss_jj(s, s, R) for s in sep])) / tp
    assert np.max(np.abs(r0 / w0 - 1)) < 2e-5
    wg = np.array([[_gauss_jj(b, a, R) for b in chi] for a in chi]) / tp
    assert np.max(np.abs(ren / wg - 1)) < 2e-5


def test_coffe_integral_n4_l0(engine):
    """Coffe.integral(r, n=4, l=0): Akima of the table at the reference's separations;
    node values exactly, and between nodes the oracle's direct integral to 1e-4."""
    from coffe_amd.coffe import Coffe
    from coffe_amd.engine import COFFE_SEP
    k, pk = _class_spectrum()
    c = Coffe()
    c.set_power_spectrum_linear(k, pk)
    print("Coffe.integral n=4 l=0", flush=True)
    oi, _, _ = O.divergent_tables(k, pk, COFFE_SEP, np.zeros(0))
    for i in (12, 40, 120, 200):
        assert abs(c.integral(r=float(COFFE_SEP[i]), n=4, l=0) / oi[i] - 1) < 1e-12
    for r in (2.5, 37.0, 151.0, 777.0):
        want, _, _ = O.divergent_tables(k, pk, np.array([r]), np.zeros(0))
        assert abs(c.integral(r=r, n=4, l=0) / want[0] - 1) < 1e-4, r
