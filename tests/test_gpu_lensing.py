"""Flat-sky lensing multipoles on the GPU (coffe_flat_lensing_multipoles and the
Coffe API), against the reference's fixtures the way its own tests hold them
(tests/test_coffe.py:328-380: rtol 5e-4) and against the oracle (oracle/lensing.py)
on the same tables (multipoles and correlation function): the kernel's 128-panel
Gauss-Legendre x integral against the oracle's 2048-panel one to 1e-9,
the closed-form density-lensing term to 1e-12, the half-integer tables of the GPU
FFTLog against the oracle's to 1e-9."""
import os

import numpy as np
import pytest

import oracle as O
from oracle import lensing as OL

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
H = 0.67
H0h = H / 2997.92458
SEPS = np.array([20.0, 40.0, 100.0, 150.0]) / H


def _pk():
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    return k * H, pk / H ** 3


def _fixture(name, l):
    x, y = np.loadtxt(os.path.join(GOLD, f"benchmark_flatsky_{name}_multipoles{l}.dat")).T
    return x, y


def test_lensing_lensing_vs_reference(engine):
    """The reference's test_multipoles_flat_lensing_lensing."""
    from coffe_amd.coffe import Coffe
    cosmo = Coffe(sep=SEPS, l=[0, 2, 4])
    cosmo.set_power_spectrum_linear(*_pk())
    cosmo.has_lensing = True
    cosmo.has_density = False
    cosmo.has_rsd = False
    cosmo.has_flatsky_nonlocal = True
    print("flat lensing-lensing via Coffe", flush=True)
    result = cosmo.compute_multipoles_bulk()
    for mp in cosmo.l:
        x, y = _fixture("lensing_lensing", mp)
        rows = [r for r in result if r.l == mp]
        assert np.allclose([r.r * H for r in rows], x)
        assert np.allclose([r.value for r in rows], y, rtol=5e-4), (mp, [r.value for r in rows], y)


def test_density_lensing_vs_reference(engine):
    """The reference's test_multipoles_flat_density_lensing (only the cross terms)."""
    from coffe_amd.coffe import Coffe
    cosmo = Coffe(sep=SEPS, l=[0, 2, 4])
    cosmo.set_power_spectrum_linear(*_pk())
    cosmo.has_lensing = True
    cosmo.has_density = True
    cosmo.has_rsd = False
    cosmo.has_flatsky_nonlocal = True
    cosmo.has_flatsky_local_nonlocal = True
    cosmo.has_only_cross_correlations = True
    print("flat density-lensing via Coffe", flush=True)
    result = cosmo.compute_multipoles_bulk()
    for mp in cosmo.l:
        x, y = _fixture("density_lensing", mp)
        rows = [r for r in result if r.l == mp]
        assert np.allclose([r.r * H for r in rows], x)
        assert np.allclose([r.value for r in rows], y, rtol=5e-4), (mp, [r.value for r in rows], y)


def test_kernel_vs_oracle_on_shared_tables(engine):
    """Both terms at 3 redshifts, l = 0..4, with magnification and galaxy biases that
    differ between the populations, on the oracle's own tables."""
    from coffe_amd import background as B
    from coffe_amd.engine import akima, flat_lensing
    k, pk = _pk()
    N = 4096
    bg = B.compute(omega_cdm=0.26, omega_baryon=0.048, h=H)
    ls = [0, 1, 2, 3, 4]
    seps = np.array([8.0, 30.0, 75.0, 160.0, 310.0])
    zs = [0.4, 1.0, 1.7]
    s1, s2, b1, b2 = 0.3, -0.1, 1.4, 0.9
    ll = {l: O.integral_table(1, l, k, pk, N) for l in (0, 2, 4)}
    dl = {(l, j): OL.half_table(l - 2 * j + 3, l - 2 * j + 1, k, pk, N, small_r=False) for l in (0, 2, 4)
          for j in range(l // 2 + 1)}
    zpar = [[z, float(akima(bg["z"], bg["chi"], [z])[0]), float(akima(bg["z"], bg["D1"], [z])[0]), s1, s2, b1, b2]
            for z in zs]
    Om = 0.26 + 0.048
    print("flat_lensing kernel, 3 z x 5 l x 5 r", flush=True)
    got_ll = flat_lensing(zpar, ls, seps, bg, ll, {}, Om, H0h, True, False)
    got_dl = flat_lensing(zpar, ls, seps, bg, {}, dl, Om, H0h, False, True)
    got_both = flat_lensing(zpar, ls, seps, bg, ll, dl, Om, H0h, True, True)
    # the sum in one launch (the last add may be fused into an FMA)
    assert np.allclose(got_both, got_ll + got_dl, rtol=1e-14, atol=0)
    for iz, z in enumerate(zs):
        for il, l in enumerate(ls):
            for ir, r in enumerate(seps):
                if l % 2:
                    assert got_ll[iz, il, ir] == 0.0 and got_dl[iz, il, ir] == 0.0
                    continue
                want_ll = OL.flat_lensing_lensing(ll[l], bg, z, l, r, s1, s2, Om, H0h)
                want_dl = OL.flat_density_lensing([dl[(l, j)] for j in range(l // 2 + 1)], bg, z, l, r, s1, s2, b1,
                                                  b2, Om, H0h)
                assert abs(got_ll[iz, il, ir] - want_ll) <= 1e-9 * abs(want_ll), (z, l, r, got_ll[iz, il, ir],
                                                                                   want_ll)
                assert abs(got_dl[iz, il, ir] - want_dl) <= 1e-12 * abs(want_dl), (z, l, r, got_dl[iz, il, ir],
                                                                                    want_dl)


def test_tables_vs_oracle(engine):
    """The GPU's half-integer tables (FFTLog with half-integer l and nu, small-r nodes by
    QAG-GK61 of J_{l+1/2}) and the (1, l) tables against the oracle's: the FFT nodes to
    1e-9, the 43 small-r nodes to 1e-6 of their largest (the GPU's quadrature runs to
    the reference's epsrel 1e-5, the oracle's to 1e-9)."""
    from coffe_amd.engine import half_integer_table, integral_table
    k, pk = _pk()
    N = 2048
    for n2, l2 in ((1, -1), (3, 1), (5, 3), (7, 5)):
        print(f"half table ({n2}/2, {l2}/2)", flush=True)
        x, y = half_integer_table(n2, l2, k, pk, N)
        xo, yo = OL.half_table(n2, l2, k, pk, N)
        assert np.allclose(x, xo, rtol=1e-14, atol=0)
        assert np.max(np.abs(y[43:] - yo[43:])) <= 1e-9 * np.max(np.abs(yo[43:])), (n2, l2)
        assert np.max(np.abs(y[:43] - yo[:43])) <= 1e-6 * np.max(np.abs(yo[:43])), (n2, l2, y[:43] - yo[:43])
    for l in (0, 2, 4):
        x, y = integral_table(1, l, k, pk, N)
        xo, yo = O.integral_table(1, l, k, pk, N)
        assert np.allclose(x, xo, rtol=1e-14, atol=0)
        assert np.max(np.abs(y - yo)) <= 1e-9 * np.max(np.abs(yo)), l


MUS = [0.0, 0.2, 0.5, 0.8, 0.95]


def test_lensing_lensing_corrfunc_vs_reference(engine):
    """tests/test_corrfunc.c:360-420 of the reference through the Python API: the
    flat-sky lensing-lensing correlation function at five mu."""
    from coffe_amd.coffe import Coffe
    x0, _ = np.loadtxt(os.path.join(GOLD, "benchmark_flatsky_lensing_lensing_corrfunc0.dat")).T
    cosmo = Coffe(sep=x0 / H, mu=MUS, has_density=False, has_lensing=True, has_flatsky_nonlocal=True)
    cosmo.set_power_spectrum_linear(*_pk())
    print("flat lensing-lensing corrfunc via Coffe", flush=True)
    result = cosmo.compute_corrfunc_bulk()
    for i, mu in enumerate(MUS):
        x, y = np.loadtxt(os.path.join(GOLD, f"benchmark_flatsky_lensing_lensing_corrfunc{i}.dat")).T
        rows = [r for r in result if r.mu == mu]
        assert np.allclose([r.r * H for r in rows], x)
        assert np.allclose([r.value for r in rows], y, rtol=5e-4), (mu, [r.value for r in rows], y)


def test_corrfunc_kernel_vs_oracle(engine):
    """Both correlation-function terms on the oracle's (1/2, -1/2) table, 2 redshifts,
    mu including 0 and +-1 (the table's r = 0 node), with unequal biases."""
    from coffe_amd import background as B
    from coffe_amd.engine import akima, flat_lensing_corrfunc
    k, pk = _pk()
    bg = B.compute(omega_cdm=0.26, omega_baryon=0.048, h=H)
    table = OL.half_table(1, -1, k, pk, 4096)
    mus = [-1.0, -0.6, 0.0, 0.3, 0.9, 1.0]
    seps = np.array([5.0, 40.0, 220.0])
    zs = [0.5, 1.3]
    s1, s2, b1, b2 = 0.3, -0.1, 1.4, 0.9
    zpar = [[z, float(akima(bg["z"], bg["chi"], [z])[0]), float(akima(bg["z"], bg["D1"], [z])[0]), s1, s2, b1, b2]
            for z in zs]
    Om = 0.26 + 0.048
    print("flat_lensing_corrfunc kernel", flush=True)
    got_ll = flat_lensing_corrfunc(zpar, mus, seps, bg, table, Om, H0h, True, False)
    got_dl = flat_lensing_corrfunc(zpar, mus, seps, bg, table, Om, H0h, False, True)
    for iz, z in enumerate(zs):
        for im, mu in enumerate(mus):
            for ir, r in enumerate(seps):
                want_ll = OL.flat_lensing_lensing_cf(table, bg, z, mu, r, s1, s2, Om, H0h)
                want_dl = OL.flat_density_lensing_cf(table, bg, z, mu, r, s1, s2, b1, b2, Om, H0h)
                assert abs(got_ll[iz, im, ir] - want_ll) <= 1e-9 * abs(want_ll), (z, mu, r)
                if mu == 0.0:
                    assert got_dl[iz, im, ir] == 0.0
                else:
                    assert abs(got_dl[iz, im, ir] - want_dl) <= 1e-12 * abs(want_dl), (z, mu, r)


def test_local_and_lensing_terms_add(engine):
    """Density with lensing, flat sky, not only the cross terms: the local multipoles
    plus density-lensing plus lensing-lensing."""
    from coffe_amd.coffe import Coffe
    kw = dict(sep=SEPS, l=[0, 2, 4], has_density=True, has_flatsky_local=True, integrals_fftlog_sampling=4096)
    full = Coffe(**kw, has_lensing=True, has_flatsky_nonlocal=True, has_flatsky_local_nonlocal=True)
    full.set_power_spectrum_linear(*_pk())
    local = Coffe(**kw)
    local.set_power_spectrum_linear(*_pk())
    ll = Coffe(**dict(kw, has_density=False), has_lensing=True, has_flatsky_nonlocal=True)
    ll.set_power_spectrum_linear(*_pk())
    dl = Coffe(**kw, has_lensing=True, has_flatsky_nonlocal=True, has_flatsky_local_nonlocal=True,
               has_only_cross_correlations=True)
    dl.set_power_spectrum_linear(*_pk())
    print("flat local + lensing sum", flush=True)
    v = lambda c: np.array([r.value for r in c.compute_multipoles_bulk()])  # noqa: E731
    assert np.allclose(v(full), v(local) + v(ll) + v(dl), rtol=1e-12, atol=0)
