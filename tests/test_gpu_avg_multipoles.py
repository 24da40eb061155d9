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


def _oracle_average(c, zmin, zmax, sep, ls, nodes=32):
    """The reference's full-sky multipoles at the node redshifts z(v) of each separation,
    weighted by h H0 / H(z) (signal.c:249-311): the oracle's binary128 weights under the
    engine's GL-256 rule (an FP64 evaluation of the reference's own form carries ~1e-7 of
    cancellation noise at ~15 Mpc and z ~ 1) times the oracle's tables."""
    import oracle as O
    bg = c._background()
    bgo = c._bg_tables()
    tabs = c._integral_tables()
    x, w = np.polynomial.legendre.leggauss(nodes)
    v, w = 0.5 * (x + 1), 0.5 * w
    chi_at = lambda z: O.akima(bg["z"], bg["chi"], np.atleast_1d(z))   # noqa: E731
    z_at = lambda ch: O.akima(bg["chi"], bg["z"], np.atleast_1d(ch))   # noqa: E731
    zb, b = np.linspace(0, 15, 16), np.ones(16)
    out = np.zeros((len(ls), len(sep)))
    for i, r in enumerate(sep):
        z1 = z_at(chi_at(zmin)[0] + r / 2)[0]
        z2 = z_at(chi_at(zmax)[0] - r / 2)[0]
        zs = z1 + (z2 - z1) * v
        hp = O.mu_weights_hp(bgo, zb, b, b, zs, ls, [r], den=c.has_density, rsd=c.has_rsd, d1=c.has_d1,
                             flatsky=False, nmu=256)[:, :, :, 0]
        It = np.array([O.akima(tabs[t][0], tabs[t][1], [r])[0] for t in O.TABLE_ORDER])
        mp = hp @ It
        out[:, i] = (w * (H / 2997.92458) / O.akima(bg["z"], bg["Hz"], zs)) @ mp
    return out


def test_average_multipoles_bin_from_zero(engine):
    """A bin starting at z = 0 with separations from 5 to 150 Mpc/h: every (z(v), r) node
    pair is its own point (paired weights), so the small separations' nodes near chi =
    r/2 are legal; against the oracle's average on the same nodes, 1e-8."""
    from coffe_amd.coffe import Coffe
    sep = np.array([5, 10, 40, 100, 150]) / H
    c = Coffe(sep=sep, l=[0, 2, 4], z_min=[0.0], z_max=[0.3], integrals_fftlog_sampling=4096)
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    c.set_power_spectrum_linear(k * H, pk / H ** 3)
    c.has_rsd = True
    print("average multipoles: bin [0, 0.3]", flush=True)
    got = c._average_values(0.0, 0.3, sep, [0, 2, 4])
    ref = _oracle_average(c, 0.0, 0.3, sep, [0, 2, 4])
    rel = np.abs(got - ref) / np.abs(ref)
    print(f"  max rel {np.max(rel):.2e}", flush=True)
    assert np.max(rel) <= 1e-8


def test_average_multipoles_wide_bin_nodes(engine):
    """A wide bin [0.5, 2.0]: the 32-node rule in the bin against 64 nodes (the quadrature
    error of the average) and against the oracle with 32 nodes (parity)."""
    from coffe_amd.coffe import Coffe
    sep = np.array([10, 40, 100, 200]) / H
    c = Coffe(sep=sep, l=[0, 2, 4], z_min=[0.5], z_max=[2.0], integrals_fftlog_sampling=4096)
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    c.set_power_spectrum_linear(k * H, pk / H ** 3)
    c.has_rsd = True
    print("average multipoles: bin [0.5, 2.0], 32 vs 64 nodes", flush=True)
    a32 = c._average_values(0.5, 2.0, sep, [0, 2, 4], nodes=32)
    a64 = c._average_values(0.5, 2.0, sep, [0, 2, 4], nodes=64)
    conv = np.max(np.abs(a32 - a64) / np.abs(a64))
    print(f"  32 vs 64 nodes: max rel {conv:.2e}", flush=True)
    assert conv <= 1e-10
    ref = _oracle_average(c, 0.5, 2.0, sep, [0, 2, 4])
    assert np.max(np.abs(a32 - ref) / np.abs(ref)) <= 1e-8
