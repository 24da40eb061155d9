"""Background tables (host C in the engine library) against the reference fixture
benchmark_background.dat (test_coffe.py:143-194 uses np.isclose with rtol 1e-5)."""
import os

import numpy as np

import oracle as O
from coffe_amd import background as B

GOLD = os.path.join(os.path.dirname(__file__), "golden")
H = 0.67


def test_background_vs_reference_fixture():
    bg = B.compute(omega_cdm=0.25, omega_baryon=0.05, h=H)
    d = np.loadtxt(os.path.join(GOLD, "benchmark_background.dat"))
    z = d[:, 0]
    ev = lambda key, zz: O.akima(bg["z"], bg[key], zz)   # noqa: E731
    assert np.allclose(ev("a", z), d[:, 1], rtol=1e-5)
    assert np.allclose(ev("Hz", z), d[:, 2] * H, rtol=1e-5)
    assert np.allclose(ev("conformal_Hz", z), d[:, 3] * H, rtol=1e-5)
    assert np.allclose(ev("conformal_Hz_prime", z), d[:, 4] * H ** 2, rtol=1e-5)
    assert np.allclose(ev("D1", z), d[:, 5] / d[0, 5], rtol=1e-5)
    assert np.allclose(ev("f", z), d[:, 6], rtol=1e-5)
    assert np.allclose(ev("chi", z), d[:, 7] / H, rtol=1e-5)
