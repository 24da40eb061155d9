"""The correlation function, the cross-correlations-only switch and the averaged
covariance through the Python API and the reference's C entry points:

- compute_corrfunc_bulk (records z, mu, r as corrfunc.c:83-90) against the oracle's
  functions_nonintegrated at each mu on the engine's own tables: flat sky to 1e-12 of
  the largest |xi| at that separation; full sky to 1e-11 of the reference's formulas
  evaluated in binary128 (the FP64 form itself is within 1e-7, its cancellation noise);
- C coffe_integrate(CORRFUNC) equals compute_corrfunc to 1e-12;
- has_only_cross_correlations: the multipoles of density + RSD + Doppler-1 without the
  auto terms equal the full ones minus each flag alone (the terms are additive), in
  the Python API and the C layer;
- compute_average_covariance_bulk equals the C layer's COVARIANCE_AVERAGE_MULTIPOLES
  (the volume integrals differ only by their quadratures, 1e-10)."""
import ctypes
import os

import numpy as np
import pytest

import oracle as O
import tests.refabi_host as R

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
H = 0.67


def _pk():
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    return k * H, pk / H ** 3


@pytest.mark.parametrize("flat", [True, False])
def test_corrfunc_vs_oracle(engine, flat):
    from coffe_amd.coffe import Coffe
    mus = [-0.95, -0.5, 0.0, 0.3, 0.7, 0.99]
    sep = np.array([40.0, 90.0, 140.0])
    c = Coffe(sep=sep, mu=mus, z_mean=[0.8, 1.5], has_density=True, has_rsd=True, has_d1=True,
              has_flatsky_local=flat, integrals_fftlog_sampling=4096)
    c.set_power_spectrum_linear(*_pk())
    zb = np.linspace(0, 15, 16)
    c.set_galaxy_bias1(zb, 1.2 + 0.3 * zb)
    c.set_galaxy_bias2(zb, 0.9 + 0.5 * zb)
    print(f"corrfunc flat={flat}", flush=True)
    res = c.compute_corrfunc_bulk()
    assert len(res) == 2 * len(mus) * len(sep)
    got = np.array([r.value for r in res]).reshape(2, len(mus), len(sep))
    assert [(r.z, r.mu, r.r) for r in res[:len(sep) + 1]] == \
        [(0.8, mus[0], s) for s in sep] + [(0.8, mus[1], sep[0])]
    bgo = c._bg_tables()
    tabs = c._integral_tables()
    b1, b2 = 1.2 + 0.3 * zb, 0.9 + 0.5 * zb
    ref = O.corrfunc(tabs, bgo, zb, b1, b2, [0.8, 1.5], mus, sep, den=True, rsd=True, d1=True, flatsky=flat)
    scale = np.max(np.abs(ref), axis=1, keepdims=True)
    # the reference's own form in FP64: exact for flat sky; full sky carries the
    # cancellation of its (chi/r)^n terms, ~1e-8 of |xi| at r = 40 and z = 1.5
    tol = 1e-12 if flat else 1e-7
    assert np.all(np.abs(got - ref) <= tol * scale), np.max(np.abs(got - ref) / scale)
    if not flat:
        # the truth: the reference's formulas in binary128 (one mu node of weight 2, l = 0)
        # times the same tables
        It = np.array([[O.akima(tabs[t][0], tabs[t][1], [r])[0] for r in sep] for t in O.TABLE_ORDER])
        hp = np.empty_like(got)
        for j, mu in enumerate(mus):
            A = O.mu_weights_hp(bgo, zb, b1, b2, [0.8, 1.5], [0], sep, den=True, rsd=True, d1=True,
                                nodes=([mu], [2.0]))[:, 0]                                  # [z][t][r]
            hp[:, j, :] = np.einsum("ztr,tr->zr", A, It)
        assert np.all(np.abs(got - hp) <= 1e-11 * scale), np.max(np.abs(got - hp) / scale)
    one = c.compute_corrfunc(z=1.5, r=90.0, mu=0.3)
    assert abs(one - got[1, 3, 1]) <= 1e-13 * abs(got[1, 3, 1])
    # the C layer's CORRFUNC point
    from coffe_amd._lib import load_library
    L = R.declare(load_library())
    host = R.Host(L)
    try:
        prog = R.Program(host, c, output_type=R.CORRFUNC, bessel_bins=4096)
        ia = R.IntegralArray()
        assert L.coffe_integrals_init(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(ia)) == 0
        v = L.coffe_integrate(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(ia), 1.5, 0.0, 0.0, 90.0,
                              0.3, 0, R.NONINTEGRATED, R.CORRFUNC)
        assert abs(v - one) <= 1e-12 * abs(one)
        L.coffe_integrals_free(ctypes.byref(ia))
    finally:
        host.close()


def test_only_cross_correlations(engine):
    from coffe_amd.coffe import Coffe
    flags = dict(has_density=True, has_rsd=True, has_d1=True)

    def mp(only_cross=False, **fl):
        c = Coffe(sep=np.array([20.0, 60.0, 150.0]), l=[0, 1, 2, 3, 4], z_mean=[1.0],
                  integrals_fftlog_sampling=4096)
        c.set_power_spectrum_linear(*_pk())
        c.reset_contributions()
        c.set_parameters(**fl)
        c.has_only_cross_correlations = only_cross
        return c, np.array([r.value for r in c.compute_multipoles_bulk()])

    c, cross = mp(True, **flags)
    _, full = mp(**flags)
    autos = sum(mp(**{k: True})[1] for k in flags)
    scale = np.max(np.abs(full))
    assert np.max(np.abs(cross - (full - autos))) <= 1e-12 * scale
    assert np.max(np.abs(cross)) > 1e-3 * scale
    c.has_d2 = True          # d2 joins as cross terms only (its auto left out; tests/test_gpu_dg.py)
    with_d2 = np.array([r.value for r in c.compute_multipoles_bulk()])
    assert np.isfinite(with_d2).all() and np.max(np.abs(with_d2 - cross)) > 0
    c.has_d2 = False
    from coffe_amd._lib import load_library
    L = R.declare(load_library())
    host = R.Host(L)
    try:
        prog = R.Program(host, c, bessel_bins=4096)
        assert prog.par.only_cross_correlations == 1
        ia = R.IntegralArray()
        assert L.coffe_integrals_init(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(ia)) == 0
        m = R.MultipolesArray()
        assert L.coffe_multipoles_init(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(ia),
                                       ctypes.byref(m)) == 0
        got = np.array([m.array[i].value for i in range(m.size)])
        assert np.max(np.abs(got - cross)) <= 1e-12 * scale
        L.coffe_multipoles_free(ctypes.byref(m))
        L.coffe_integrals_free(ctypes.byref(ia))
    finally:
        host.close()


def test_average_covariance_python_vs_c(engine):
    from coffe_amd.coffe import Coffe
    c = Coffe(has_density=True, has_rsd=True, number_density1=[1e-3 * H ** 3, 5e-4 * H ** 3],
              number_density2=[1e-3 * H ** 3, 5e-4 * H ** 3], z_min=[0.8, 1.2], z_max=[1.0, 1.6], deltaz=[0.1, 0.2],
              fsky=[0.2, 0.3], pixelsize=[50 / H, 40 / H], sep=np.arange(50, 350, 50) / H, l=[0, 2, 4],
              covariance_integration_sampling=2048)
    c.set_power_spectrum_linear(*_pk())
    print("average covariance", flush=True)
    py = c.compute_average_covariance_bulk()
    assert len(py) == 2 * 18 * 18
    from coffe_amd._lib import load_library
    L = R.declare(load_library())
    host = R.Host(L)
    try:
        prog = R.Program(host, c, output_type=R.COVARIANCE_AVERAGE_MULTIPOLES)
        cm, ca = R.CovarianceArray(), R.AverageCovarianceArray()
        assert L.coffe_covariance_init(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(cm),
                                       ctypes.byref(ca)) == 0
        assert ca.size == len(py)
        for i, b in enumerate(py):
            a = ca.array[i]
            assert (a.coords.z_min, a.coords.z_max, a.coords.l1, a.coords.l2, a.coords.separation1,
                    a.coords.separation2) == (b.z_min, b.z_max, b.l1, b.l2, b.r1, b.r2)
            assert abs(a.value - b.value) <= 1e-10 * abs(b.value) + 1e-300
        L.coffe_average_covariance_free(ctypes.byref(ca))
    finally:
        host.close()
    M = np.array([r.value for r in py[:18 * 18]]).reshape(18, 18)
    assert np.allclose(M, M.T, rtol=1e-6, atol=1e-6 * np.max(np.abs(M)))
