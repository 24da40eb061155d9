"""The drop-in Coffe object on the GPU, driven exactly as the reference's own Python
tests drive COFFE (test_coffe.py:196-233, 270-303, 382-430) against their fixtures:
integrals rtol 1e-4, multipoles rtol 5e-4, covariance (2D FFTLog here, the direct
quadrature there) 2e-3 of each block's largest entry, symmetric and positive definite."""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
H = 0.67
COFFE_H0 = 1 / 2997.92458


def _pk():
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    return k * H, pk / H ** 3


def test_integrals(engine):
    from coffe_amd.coffe import Coffe
    cosmo = Coffe()
    cosmo.set_power_spectrum_linear(*_pk())
    mapping = [(0, 0), (0, 2), (0, 4), (1, 1), (1, 3), (2, 0), (2, 2), (3, 1)]
    for index, (n, l) in enumerate(mapping):
        xarr, yarr = np.loadtxt(os.path.join(GOLD, f"benchmark_integral{index}.dat")).T
        for x, y in zip(xarr, yarr):
            if 1.5 < x / (COFFE_H0 * H) < 20000:
                result = cosmo.integral(r=x / (COFFE_H0 * H), l=l, n=n)
                if n > l:
                    result *= (COFFE_H0 * H) ** (n - l)
                assert np.isclose(y, result, rtol=1e-4), (n, l, x)
    # below the FFT grid the small-separation nodes are used
    assert np.isfinite(cosmo.integral(r=0.0, n=0, l=0)) and cosmo.integral(r=0.0, n=0, l=2) == 0.0
    assert np.isfinite(cosmo.integral(r=0.5, n=1, l=1))


@pytest.mark.parametrize("prefix,flags", [("den", dict(has_density=True)), ("rsd", dict(has_rsd=True)),
                                          ("d1", dict(has_d1=True)),
                                          ("std", dict(has_density=True, has_rsd=True)),
                                          ("std_d1", dict(has_density=True, has_rsd=True, has_d1=True))])
def test_multipoles(engine, prefix, flags):
    from coffe_amd.coffe import Coffe
    cosmo = Coffe(sep=np.array([10, 20, 40, 100, 150]) / H, l=[0, 2, 4])
    cosmo.set_power_spectrum_linear(*_pk())
    cosmo.reset_contributions()
    cosmo.set_parameters(**flags)
    print(f"Coffe multipoles {prefix}", flush=True)
    result = cosmo.compute_multipoles_bulk()
    assert len(result) == 15
    for mp in cosmo.l:
        x, y = np.loadtxt(os.path.join(GOLD, f"benchmark_{prefix}_multipoles{mp}.dat")).T
        rows = [r for r in result if r.l == mp]
        assert np.allclose([r.r * H for r in rows], x)
        assert np.allclose([r.value for r in rows], y, rtol=5e-4), (prefix, mp)
    single = cosmo.compute_multipole(z=1.0, r=float(40 / H), l=2)
    bulk = [r.value for r in result if r.l == 2 and abs(r.r - 40 / H) < 1e-9][0]
    assert abs(single - bulk) <= 1e-12 * abs(bulk)


def test_covariance_multipoles(engine):
    from coffe_amd.coffe import Coffe
    cosmo = Coffe()
    cosmo.set_parameters(has_density=True, has_rsd=True, number_density1=[1e-3 * H ** 3],
                         number_density2=[1e-3 * H ** 3], z_mean=[1.0], deltaz=[0.1], fsky=[0.2],
                         pixelsize=[50 / H], sep=np.arange(50, 350, 50) / H, l=[0, 2, 4])
    cosmo.set_power_spectrum_linear(*_pk())
    print("Coffe covariance", flush=True)
    result = cosmo.compute_covariance_bulk()
    R = 6
    for mp1 in cosmo.l:
        for mp2 in cosmo.l:
            x, y, z = np.loadtxt(os.path.join(GOLD, f"benchmark_multipoles_covariance_{mp1}{mp2}.dat")).T
            rows = [r for r in result if r.l1 == mp1 and r.l2 == mp2]
            assert np.allclose([r.r1 * H for r in rows], x)
            assert np.allclose([r.r2 * H for r in rows], y)
            v = np.array([r.value for r in rows])
            assert np.max(np.abs(v - z)) <= 2e-3 * np.max(np.abs(z)), (mp1, mp2)
    M = np.array([r.value for r in result]).reshape(3 * R, 3 * R)
    assert np.allclose(M, M.T, rtol=1e-6, atol=1e-6 * np.max(np.abs(M)))
    assert np.all(np.linalg.eigvalsh(0.5 * (M + M.T)) > 0)


def test_covariance_populations_api(engine):
    """The Coffe object's multi-tracer covariance: with the default populations setting
    galaxy_bias2 changes nothing (the reference reads galaxy_bias1 in every slot,
    covariance.c:60-74, parser.c:945-948); populations (1, 2, 1, 2) with number_density2
    read bias 2 and density2 and equal the oracle's assembly of the same integrals."""
    import torch
    from coffe_amd.coffe import Coffe
    from coffe_amd.covariance import CovPlan
    zb = np.linspace(0, 3, 16)
    common = dict(has_density=True, has_rsd=True, number_density1=[1e-3 * H ** 3],
                  number_density2=[3e-4 * H ** 3], z_mean=[1.0], deltaz=[0.1], fsky=[0.2], pixelsize=[50 / H],
                  sep=np.arange(50, 350, 50) / H, l=[0, 2, 4], covariance_integration_sampling=1024)
    base = Coffe(**common)
    base.set_power_spectrum_linear(*_pk())
    base.set_galaxy_bias1(zb, 1.4 * np.ones_like(zb))
    v0 = np.array([r.value for r in base.compute_covariance_bulk()])
    base.set_galaxy_bias2(zb, 2.3 * np.ones_like(zb))
    v1 = np.array([r.value for r in base.compute_covariance_bulk()])
    assert np.array_equal(v0, v1)
    print("Coffe covariance populations (1, 2, 1, 2)", flush=True)
    base.covariance_populations = [1, 2, 1, 2]
    v2 = np.array([r.value for r in base.compute_covariance_bulk()])
    # the same integrals through the oracle's assembly
    k, pk = _pk()
    sep = base.sep
    plan = CovPlan(base._engine_handle(), 1024, k[0], k[-1], [0, 2, 4], sep)
    integ = plan.integrals(torch.tensor(k, device="cuda"), torch.tensor(pk, device="cuda"), [len(k)]).cpu().numpy()[0]
    plan.close()
    pk_int = np.zeros((9, 36))
    pk2_int = np.zeros((9, 36))
    for q, (i, j) in enumerate([(i, j) for i in range(3) for j in range(i, 3)]):
        pk_int[i * 3 + j] = integ[0, q].reshape(-1)
        pk2_int[i * 3 + j] = integ[1, q].reshape(-1)
    z, dz, fsky = 1.0, 0.1, 0.2
    vol = 4 * np.pi * fsky * (base.comoving_distance(z + dz) ** 3 - base.comoving_distance(z - dz) ** 3) / 3
    ref = O.covariance_assemble([0, 2, 4], sep, np.array([[1.4, 2.3, 1.0, 1.0]]), [base.growth_rate(z)],
                                [base.growth_factor(z)], [1e-3 * H ** 3], [vol], [50 / H], pk_int, pk2_int,
                                pops=(1, 2, 1, 2), nbar2=[3e-4 * H ** 3]).reshape(-1)
    assert np.max(np.abs(v2 - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert np.max(np.abs(v2 - v1)) > 1e-3 * np.max(np.abs(v1))


def test_plain_configuration_end_to_end(engine):
    """The plain configuration through the Coffe object: flat-sky density + RSD at z = 1,
    b = 1.5, LCDM Omega_m = 0.31 (Omega_cdm 0.26, Omega_b 0.05, h 0.67); the synthetic
    P(k) on 4096 log-spaced samples in [1e-5, 10] h/Mpc; FFTLog with N = 10000; l = 0, 2, 4
    at 64 separations 10-200 Mpc/h.  Against the oracle's reference sequence on the same
    background tables (the 8 tables with their small-r nodes, Akima lookups, QAG-GK61 mu
    integral), 1e-8 relative per point (1e-6 of the multipole's largest |value| as floor)."""
    from coffe_amd.coffe import Coffe
    from tests.synth import spectrum
    sep = np.linspace(10.0, 200.0, 64) / H
    cosmo = Coffe(omega_cdm=0.26, omega_baryon=0.05, h=H, has_density=True, has_rsd=True, has_flatsky_local=True,
                  z_mean=[1.0], l=[0, 2, 4], sep=sep, integrals_fftlog_sampling=10000)
    k, p = spectrum(4096)
    cosmo.set_power_spectrum_linear(k, p)
    zb = np.linspace(0.0, 3.0, 16)
    cosmo.set_galaxy_bias1(zb, 1.5 * np.ones_like(zb))
    cosmo.set_galaxy_bias2(zb, 1.5 * np.ones_like(zb))
    print("plain configuration: Coffe", flush=True)
    got = np.array([r.value for r in cosmo.compute_multipoles_bulk()]).reshape(3, 64)
    print("plain configuration: oracle", flush=True)
    ref = O.reference_batch([(k, p)], [10000], cosmo._bg_tables(), zb, 1.5 * np.ones_like(zb), 1.5 * np.ones_like(zb),
                            [1.0], [0, 2, 4], sep, den=True, rsd=True, d1=False, flatsky=True)[0, 0]
    for il in range(3):
        floor = 1e-6 * np.max(np.abs(ref[il]))
        rel = np.abs(got[il] - ref[il]) / np.maximum(np.abs(ref[il]), floor)
        print(f"  l={2 * il}: max rel {np.max(rel):.2e}", flush=True)
        assert np.max(rel) <= 1e-8, (il, np.max(rel))
