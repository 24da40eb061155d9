"""Redshift-averaged multipoles through the Python API (Coffe.compute_average_multipoles_bulk:
the engine's GPU mu weights at the (z, r) node pairs, Gauss-Legendre in the bin) in the
configuration of the reference's test (test_coffe.py:305-326): density + RSD, bin
[0.9, 1.1], 5 separations, against benchmark_avg_multipoles{0,2,4}.dat (the reference's
rtol 5e-4; 5e-5 here)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
H = 0.67


def test_average_multipoles_fixture(engine):
    from coffe_amd.coffe import Coffe
    c = Coffe(sep=np.array([10, 20, 40, 100, 150]) / H, l=[0, 2, 4], z_min=[0.9], z_max=[1.1])
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    c.set_power_spectrum_linear(k * H, pk / H ** 3)
    c.has_density = True
    c.has_rsd = True
    print("average multipoles: bin [0.9, 1.1], 5 seps", flush=True)
    res = c.compute_average_multipoles_bulk()
    assert len(res) == 15
    for l in (0, 2, 4):
        x, y = np.loadtxt(os.path.join(GOLD, f"benchmark_avg_multipoles{l}.dat")).T
        pts = [p for p in res if p.l == l]
        assert all(p.z_min == 0.9 and p.z_max == 1.1 for p in pts)
        assert np.allclose(np.array([p.r for p in pts]) * H, x)
        got = np.array([p.value for p in pts])
        assert np.allclose(got, y, rtol=5e-4)
        assert np.max(np.abs(got / y - 1)) < 5e-5, (l, got / y - 1)


def test_average_multipoles_chunking(engine):
    """The weights computed two separations at a time equal the single-chunk ones
    (the matching (node, r) pairs are the same wavefronts' work)."""
    from coffe_amd.coffe import Coffe
    sep = np.array([10, 20, 40, 100, 150]) / H
    c = Coffe(sep=sep, l=[0, 2, 4])
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    c.set_power_spectrum_linear(k * H, pk / H ** 3)
    c.has_rsd = True
    print("average multipoles: chunks of 2 vs 8", flush=True)
    a = c._average_values(0.9, 1.1, sep, [0, 2, 4], chunk=8)
    b = c._average_values(0.9, 1.1, sep, [0, 2, 4], chunk=2)
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(a))
