"""The reference's C entry points (include/coffe_refabi.h) driven as a COFFE C program
drives them (main.c: coffe_integrals_init, then coffe_multipoles_init /
coffe_average_multipoles_init / coffe_covariance_init, coffe_integrate, the free
routines), with the host's structures filled from a configured Coffe object
(tests/refabi_host.py):

- the reference's fixtures at the Python API's tolerances (integrals rtol 1e-4,
  multipoles 5e-4, averaged multipoles 5e-5, covariance 2e-3 of each block's maximum);
- the Python API (same kernels, same inputs) to 1e-12: the drop-in adds only its host
  plumbing;
- coffe_integrate's one-point outputs: MULTIPOLES equals the batch, CORRFUNC equals
  sum_l M_l P_l(mu) (exact for flat-sky density + RSD, whose only multipoles are l = 0,
  2, 4), AVERAGE_MULTIPOLES equals the averaged batch;
- the averaged covariance against the same integrals assembled with the volume of
  covariance.c:808-837 from scipy's quadrature."""
import ctypes
import os

import numpy as np
import pytest

import tests.refabi_host as R

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
H = 0.67
TABLES = [(0, 0), (0, 2), (0, 4), (1, 1), (1, 3), (2, 0), (2, 2), (3, 1)]


def _pk():
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    return k * H, pk / H ** 3


@pytest.fixture
def lib(engine):
    from coffe_amd._lib import load_library
    L = R.declare(load_library())
    host = R.Host(L)
    yield L, host
    host.close()


def _integrals(L, prog):
    ia = R.IntegralArray()
    rc = L.coffe_integrals_init(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(ia))
    assert rc == 0, rc
    return ia


def test_integrals_init(lib):
    """The eight tables are the engine's full tables (N + 43 nodes) in host splines, equal
    to the reference's fixtures; with `divergent` the I_0^4 table, its zero-separation
    renormalisation and the 201 x 201 grid on [0, chi(max(z_mean + deltaz))]."""
    from coffe_amd.coffe import Coffe
    from coffe_amd.engine import COFFE_SEP, divergent_tables, integral_table
    L, host = lib
    cosmo = Coffe(z_mean=[1.0], deltaz=[0.2])
    k, pk = _pk()
    cosmo.set_power_spectrum_linear(k, pk)
    prog = R.Program(host, cosmo, divergent=1)
    print("coffe_integrals_init, divergent", flush=True)
    ia = _integrals(L, prog)
    assert ia.size == 9
    for idx, (n, l) in enumerate(TABLES):
        I = L.coffe_find_integral(ctypes.byref(ia), n, l, 0, 0).contents
        x, y = host.nodes(I.result)
        ex, ey = integral_table(n, l, k, pk, 10000)
        assert np.array_equal(x, ex) and np.array_equal(y, ey), (n, l)
        fx, fy = np.loadtxt(os.path.join(GOLD, f"benchmark_integral{idx}.dat")).T
        sel = (fx / (R.COFFE_H0 * H) > 1.5) & (fx / (R.COFFE_H0 * H) < 20000)
        got = np.array([host._eval(ctypes.pointer(I.result), v) for v in fx[sel] / (R.COFFE_H0 * H)])
        if n > l:
            got *= (R.COFFE_H0 * H) ** (n - l)
        assert np.allclose(got, fy[sel], rtol=1e-4), (n, l)
    I40 = L.coffe_find_integral(ctypes.byref(ia), 4, 0, 0, 0).contents
    # (integrals that reach the interval cap keep their values, as the reference's do)
    sep, i40, r0, _, _, _ = divergent_tables(k, pk, return_status=True)
    x, y = host.nodes(I40.result)
    assert np.array_equal(x, COFFE_SEP) and np.array_equal(y, i40)
    x0, y0 = host.nodes(I40.renormalization_zero_separation)
    assert x0[0] == 0.0 and y0[0] == 0.0 and np.array_equal(x0[1:], COFFE_SEP[1:]) and np.array_equal(y0[1:], r0[1:])
    gx, gy, gz = host.grid(I40.renormalization)
    chi, ren = cosmo._ren_grid(cosmo.comoving_distance(1.2))
    assert np.array_equal(gx, chi) and np.array_equal(gy, chi)
    assert np.allclose(gz, ren.reshape(-1), rtol=1e-13, atol=0)
    before = len(host.objects)
    assert L.coffe_integrals_free(ctypes.byref(ia)) == 0 and ia.size == 0
    assert before - len(host.objects) == 11   # 8 + I_0^4, its zero-separation spline, the grid


@pytest.mark.parametrize("prefix,flags", [("den", dict(has_density=True)), ("rsd", dict(has_rsd=True)),
                                          ("d1", dict(has_d1=True)),
                                          ("std", dict(has_density=True, has_rsd=True)),
                                          ("std_d1", dict(has_density=True, has_rsd=True, has_d1=True)),
                                          ("d2", dict(has_d2=True)), ("g1", dict(has_g1=True))])
def test_multipoles_init(lib, prefix, flags):
    from coffe_amd.coffe import Coffe
    L, host = lib
    cosmo = Coffe(sep=np.array([10, 20, 40, 100, 150]) / H, l=[0, 2, 4])
    cosmo.set_power_spectrum_linear(*_pk())
    cosmo.reset_contributions()
    cosmo.set_parameters(**flags)
    prog = R.Program(host, cosmo)
    print(f"coffe_multipoles_init {prefix}", flush=True)
    ia = _integrals(L, prog)
    mp = R.MultipolesArray()
    assert L.coffe_multipoles_init(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(ia),
                                   ctypes.byref(mp)) == 0
    assert mp.size == 15
    recs = [mp.array[i] for i in range(mp.size)]
    py = cosmo.compute_multipoles_bulk()
    for a, b in zip(recs, py):
        assert (a.coords.z_mean, a.coords.separation, a.coords.l) == (b.z, b.r, b.l)
        assert abs(a.value - b.value) <= 1e-12 * abs(b.value), (prefix, b.l, b.r, a.value, b.value)
    # the reference's own criterion (test_coffe.py: np.allclose, rtol 5e-4): its FP64
    # full-sky forms carry cancellation noise where a multipole is ~1e-7 of the monopole
    # (the d1 hexadecapole), which the Python API comparison above does not share
    for mpl in (0, 2, 4):
        x, y = np.loadtxt(os.path.join(GOLD, f"benchmark_{prefix}_multipoles{mpl}.dat")).T
        got = np.array([r.value for r in recs if r.coords.l == mpl])
        assert np.allclose(got, y, rtol=5e-4), (prefix, mpl)
    # one point through coffe_integrate equals the batch
    rec = recs[7]
    one = L.coffe_integrate(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(ia), rec.coords.z_mean, 0.0,
                            0.0, rec.coords.separation, 0.0, rec.coords.l, R.NONINTEGRATED, R.MULTIPOLES)
    assert abs(one - rec.value) <= 1e-12 * abs(rec.value)
    assert L.coffe_multipoles_free(ctypes.byref(mp)) == 0 and mp.size == 0
    assert L.coffe_integrals_free(ctypes.byref(ia)) == 0


def test_integrate_corrfunc_is_legendre_sum(lib):
    """CORRFUNC at mu (the local correlation function, one mu node) equals sum_l M_l(r)
    P_l(mu) over l = 0, 2, 4 for flat-sky density + RSD (no other multipole is non-zero)."""
    from numpy.polynomial import legendre
    from coffe_amd.coffe import Coffe
    L, host = lib
    cosmo = Coffe(sep=np.array([20.0, 60.0, 130.0]) / H, l=[0, 2, 4], has_density=True, has_rsd=True,
                  has_flatsky_local=True, integrals_fftlog_sampling=4096)
    cosmo.set_power_spectrum_linear(*_pk())
    prog = R.Program(host, cosmo)
    ia = _integrals(L, prog)
    mp = R.MultipolesArray()
    assert L.coffe_multipoles_init(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(ia),
                                   ctypes.byref(mp)) == 0
    M = np.array([mp.array[i].value for i in range(mp.size)]).reshape(3, 3)   # [l][r]
    for ir, r in enumerate(cosmo.sep):
        for mu in (-0.9, -0.3, 0.0, 0.45, 1.0):
            xi = L.coffe_integrate(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(ia), 1.0, 0.0, 0.0,
                                   float(r), mu, 0, R.NONINTEGRATED, R.CORRFUNC)
            ref = sum(M[j, ir] * legendre.legval(mu, [0] * l + [1]) for j, l in enumerate((0, 2, 4)))
            assert abs(xi - ref) <= 1e-11 * np.max(np.abs(M[:, ir])), (r, mu, xi, ref)
    L.coffe_multipoles_free(ctypes.byref(mp))
    L.coffe_integrals_free(ctypes.byref(ia))


def test_average_multipoles_init(lib):
    from coffe_amd.coffe import Coffe
    L, host = lib
    cosmo = Coffe(sep=np.array([10, 20, 40, 100, 150]) / H, l=[0, 2, 4], z_min=[0.9], z_max=[1.1],
                  has_density=True, has_rsd=True)
    cosmo.set_power_spectrum_linear(*_pk())
    prog = R.Program(host, cosmo, output_type=R.AVERAGE_MULTIPOLES)
    print("coffe_average_multipoles_init", flush=True)
    ia = _integrals(L, prog)
    ramp = R.AverageMultipolesArray()
    assert L.coffe_average_multipoles_init(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(ia),
                                           ctypes.byref(ramp)) == 0
    recs = [ramp.array[i] for i in range(ramp.size)]
    py = cosmo.compute_average_multipoles_bulk()
    for a, b in zip(recs, py):
        assert (a.coords.z_min, a.coords.z_max, a.coords.separation, a.coords.l) == (b.z_min, b.z_max, b.r, b.l)
        assert abs(a.value - b.value) <= 1e-12 * abs(b.value)
    for mpl in (0, 2, 4):
        x, y = np.loadtxt(os.path.join(GOLD, f"benchmark_avg_multipoles{mpl}.dat")).T
        got = np.array([r.value for r in recs if r.coords.l == mpl])
        assert np.max(np.abs(got / y - 1)) < 5e-5, mpl
    rec = recs[4]
    one = L.coffe_integrate(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(ia), 1.0, 0.9, 1.1,
                            rec.coords.separation, 0.0, rec.coords.l, R.NONINTEGRATED, R.AVERAGE_MULTIPOLES)
    assert abs(one - rec.value) <= 1e-12 * abs(rec.value)
    assert L.coffe_average_multipoles_free(ctypes.byref(ramp)) == 0
    L.coffe_integrals_free(ctypes.byref(ia))


def _covariance_cosmo(**kw):
    from coffe_amd.coffe import Coffe
    cosmo = Coffe(has_density=True, has_rsd=True, number_density1=[1e-3 * H ** 3], number_density2=[1e-3 * H ** 3],
                  z_mean=[1.0], deltaz=[0.1], fsky=[0.2], pixelsize=[50 / H], sep=np.arange(50, 350, 50) / H,
                  l=[0, 2, 4], **kw)
    cosmo.set_power_spectrum_linear(*_pk())
    return cosmo


def test_covariance_init(lib):
    L, host = lib
    cosmo = _covariance_cosmo()
    prog = R.Program(host, cosmo, output_type=R.COVARIANCE_MULTIPOLES)
    print("coffe_covariance_init", flush=True)
    cm, ca = R.CovarianceArray(), R.AverageCovarianceArray()
    assert L.coffe_covariance_init(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(cm),
                                   ctypes.byref(ca)) == 0
    assert cm.size == 3 * 6 * 3 * 6 and ca.size == 0
    recs = [cm.array[i] for i in range(cm.size)]
    py = cosmo.compute_covariance_bulk()
    for a, b in zip(recs, py):
        assert (a.coords.l1, a.coords.l2, a.coords.separation1, a.coords.separation2, a.coords.z_mean) == \
            (b.l1, b.l2, b.r1, b.r2, b.z)
        assert abs(a.value - b.value) <= 1e-12 * abs(b.value) + 1e-300
    for mp1 in (0, 2, 4):
        for mp2 in (0, 2, 4):
            _, _, z = np.loadtxt(os.path.join(GOLD, f"benchmark_multipoles_covariance_{mp1}{mp2}.dat")).T
            v = np.array([r.value for r in recs if r.coords.l1 == mp1 and r.coords.l2 == mp2])
            assert np.max(np.abs(v - z)) <= 2e-3 * np.max(np.abs(z)), (mp1, mp2)
    hit = L.coffe_covariance_find(ctypes.byref(cm), 1.0, 2, 4, 100 / H, 250 / H)
    ref = [r.value for r in recs if (r.coords.l1, r.coords.l2) == (2, 4) and
           abs(r.coords.separation1 - 100 / H) < 1e-9 and abs(r.coords.separation2 - 250 / H) < 1e-9][0]
    assert hit.value == ref
    assert L.coffe_covariance_free(ctypes.byref(cm)) == 0 and cm.size == 0


def test_average_covariance_init(lib):
    """COVARIANCE_AVERAGE_MULTIPOLES: the flat-sky covariance at z_mean = (zmin + zmax)/2
    with the volume 4 pi fsky / (int dz / (H chi^2 (1 + z)) (H0 h)^2) (covariance.c:
    808-837): equals the multipole covariance of the same integrals scaled by the ratio of
    the volumes (every term goes as 1/V), the volume integral from scipy's quad."""
    from scipy.integrate import quad
    from coffe_amd.coffe import _Spline
    L, host = lib
    zmin, zmax = 0.9, 1.1
    cosmo = _covariance_cosmo(z_min=[zmin], z_max=[zmax])
    # the multipole covariance at the bin's mid-point (z_mean = 1.0, deltaz = 0.1 -> the
    # shell volume of [0.9, 1.1])
    prog_mp = R.Program(host, cosmo, output_type=R.COVARIANCE_MULTIPOLES)
    cm, ca = R.CovarianceArray(), R.AverageCovarianceArray()
    assert L.coffe_covariance_init(ctypes.byref(prog_mp.par), ctypes.byref(prog_mp.bg), ctypes.byref(cm),
                                   ctypes.byref(ca)) == 0
    prog = R.Program(host, cosmo, output_type=R.COVARIANCE_AVERAGE_MULTIPOLES)
    cm2, ca2 = R.CovarianceArray(), R.AverageCovarianceArray()
    print("coffe_covariance_init, averaged", flush=True)
    assert L.coffe_covariance_init(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(cm2),
                                   ctypes.byref(ca2)) == 0
    assert cm2.size == 0 and ca2.size == cm.size
    bg = cosmo._background()
    chi, cH = _Spline(bg["z"], bg["chi"]), _Spline(bg["z"], bg["conformal_Hz"])
    integ = quad(lambda z: 1 / cH(z)[0] / chi(z)[0] ** 2 / (1 + z), zmin, zmax, epsabs=0, epsrel=1e-13)[0]
    hH0 = H * R.COFFE_H0
    v_avg = 4 * np.pi * 0.2 / integ / hH0 ** 2
    v_mp = 4 * np.pi * 0.2 * (chi(1.1)[0] ** 3 - chi(0.9)[0] ** 3) / 3
    for i in range(cm.size):
        a, b = ca2.array[i], cm.array[i]
        assert (a.coords.z_min, a.coords.z_max, a.coords.l1, a.coords.l2) == (zmin, zmax, b.coords.l1, b.coords.l2)
        assert abs(a.value - b.value * v_mp / v_avg) <= 1e-10 * abs(b.value * v_mp / v_avg) + 1e-300
    L.coffe_covariance_free(ctypes.byref(cm))
    L.coffe_average_covariance_free(ctypes.byref(ca2))


@pytest.mark.parametrize("name,flags", [
    ("lensing_lensing", dict(has_lensing=True, has_flatsky_nonlocal=True)),
    ("density_lensing", dict(has_density=True, has_lensing=True, has_flatsky_nonlocal=True,
                             has_flatsky_local_nonlocal=True, has_only_cross_correlations=True)),
])
def test_flat_lensing_multipoles_init(lib, name, flags):
    """The reference's flat-sky lensing runs (test_multipoles.c:393-501) through
    coffe_integrals_init (which adds the (1, l), the half-integer and the (1/2, -1/2)
    tables to the host's splines) and coffe_multipoles_init: the Python API to 1e-12, the
    reference's fixtures to 5e-4; one point of coffe_integrate (SINGLE_INTEGRATED for
    density-lensing, DOUBLE_INTEGRATED for lensing-lensing) equals the batch."""
    from coffe_amd.coffe import Coffe
    L, host = lib
    cosmo = Coffe(sep=np.array([20.0, 40.0, 100.0, 150.0]) / H, l=[0, 2, 4])
    cosmo.set_power_spectrum_linear(*_pk())
    cosmo.reset_contributions()
    cosmo.set_parameters(**flags)
    prog = R.Program(host, cosmo)
    print(f"coffe_multipoles_init flat {name}", flush=True)
    ia = _integrals(L, prog)
    mp = R.MultipolesArray()
    assert L.coffe_multipoles_init(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(ia),
                                   ctypes.byref(mp)) == 0
    recs = [mp.array[i] for i in range(mp.size)]
    py = cosmo.compute_multipoles_bulk()
    assert len(recs) == len(py) == 12
    for a, b in zip(recs, py):
        assert (a.coords.z_mean, a.coords.separation, a.coords.l) == (b.z, b.r, b.l)
        assert abs(a.value - b.value) <= 1e-12 * abs(b.value), (name, b.l, b.r, a.value, b.value)
    for mpl in (0, 2, 4):
        x, y = np.loadtxt(os.path.join(GOLD, f"benchmark_flatsky_{name}_multipoles{mpl}.dat")).T
        got = np.array([r.value for r in recs if r.coords.l == mpl])
        assert np.allclose(got, y, rtol=5e-4), (name, mpl)
    kind = R.DOUBLE_INTEGRATED if name == "lensing_lensing" else R.SINGLE_INTEGRATED
    for rec in (recs[1], recs[6], recs[11]):
        one = L.coffe_integrate(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(ia), rec.coords.z_mean,
                                0.0, 0.0, rec.coords.separation, 0.0, rec.coords.l, kind, R.MULTIPOLES)
        assert abs(one - rec.value) <= 1e-12 * abs(rec.value), (name, rec.coords.l, one, rec.value)
    # the lensing terms are not in the other integration kind
    other = R.SINGLE_INTEGRATED if kind == R.DOUBLE_INTEGRATED else R.DOUBLE_INTEGRATED
    rec = recs[1]
    assert L.coffe_integrate(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(ia), rec.coords.z_mean, 0.0,
                             0.0, rec.coords.separation, 0.0, rec.coords.l, other, R.MULTIPOLES) == 0.0
    L.coffe_multipoles_free(ctypes.byref(mp))
    L.coffe_integrals_free(ctypes.byref(ia))


def test_flat_lensing_corrfunc_integrate(lib):
    """The reference's flat-sky lensing-lensing correlation function test
    (test_corrfunc.c:360-420): NONINTEGRATED + SINGLE_INTEGRATED + DOUBLE_INTEGRATED of
    coffe_integrate at five mu against its fixtures (5e-4) and the Python API (1e-12)."""
    from coffe_amd.coffe import Coffe
    L, host = lib
    mus = [0.0, 0.2, 0.5, 0.8, 0.95]
    x0, _ = np.loadtxt(os.path.join(GOLD, "benchmark_flatsky_lensing_lensing_corrfunc0.dat")).T
    cosmo = Coffe(sep=x0 / H, mu=mus, has_density=False, has_lensing=True, has_flatsky_nonlocal=True)
    cosmo.set_power_spectrum_linear(*_pk())
    prog = R.Program(host, cosmo, output_type=R.CORRFUNC)
    print("coffe_integrate flat lensing corrfunc", flush=True)
    ia = _integrals(L, prog)
    py = {(round(r.mu, 6), round(r.r, 9)): r.value for r in cosmo.compute_corrfunc_bulk()}
    for i, mu in enumerate(mus):
        x, y = np.loadtxt(os.path.join(GOLD, f"benchmark_flatsky_lensing_lensing_corrfunc{i}.dat")).T
        got = []
        for r in x / H:
            got.append(sum(L.coffe_integrate(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(ia), 1.0,
                                             0.0, 0.0, float(r), mu, 0, kind, R.CORRFUNC)
                           for kind in (R.NONINTEGRATED, R.SINGLE_INTEGRATED, R.DOUBLE_INTEGRATED)))
            want = py[(round(mu, 6), round(float(r), 9))]
            assert abs(got[-1] - want) <= 1e-12 * abs(want), (mu, r, got[-1], want)
        assert np.allclose(got, y, rtol=5e-4), (mu, got, y)
    L.coffe_integrals_free(ctypes.byref(ia))


@pytest.mark.parametrize("flags", [
    dict(has_lensing=True),
    dict(has_g4=True, has_g5=True),
    dict(has_density=True, has_rsd=True, has_lensing=True, has_g4=True),
])
def test_fullsky_integrated_multipoles_init(lib, flags):
    """Full-sky lensing, g4 and g5 (signal.c:556-771) through coffe_integrals_init (the
    divergent table and its renormalisations for g4 / g5) and coffe_multipoles_init:
    the Python API to 1e-10 (the same kernel and mu rule); NONINTEGRATED +
    SINGLE_INTEGRATED + DOUBLE_INTEGRATED of coffe_integrate equal each record."""
    from coffe_amd.coffe import Coffe
    L, host = lib
    cosmo = Coffe(sep=np.array([20.0, 100.0]) / H, l=[0, 2])
    cosmo.set_power_spectrum_linear(*_pk())
    cosmo.reset_contributions()
    cosmo.set_parameters(**flags)
    prog = R.Program(host, cosmo)
    print("coffe_multipoles_init full sky", flags, flush=True)
    ia = _integrals(L, prog)
    mp = R.MultipolesArray()
    assert L.coffe_multipoles_init(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(ia),
                                   ctypes.byref(mp)) == 0
    recs = [mp.array[i] for i in range(mp.size)]
    py = cosmo.compute_multipoles_bulk()
    assert len(recs) == len(py) == 4
    for a, b in zip(recs, py):
        assert (a.coords.z_mean, a.coords.separation, a.coords.l) == (b.z, b.r, b.l)
        assert abs(a.value - b.value) <= 1e-10 * abs(b.value), (flags, b.l, b.r, a.value, b.value)
    for rec in (recs[0], recs[3]):
        one = sum(L.coffe_integrate(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(ia),
                                    rec.coords.z_mean, 0.0, 0.0, rec.coords.separation, 0.0, rec.coords.l, kind,
                                    R.MULTIPOLES)
                  for kind in (R.NONINTEGRATED, R.SINGLE_INTEGRATED, R.DOUBLE_INTEGRATED))
        assert abs(one - rec.value) <= 1e-10 * abs(rec.value), (flags, rec.coords.l, one, rec.value)
    L.coffe_multipoles_free(ctypes.byref(mp))
    L.coffe_integrals_free(ctypes.byref(ia))


def test_fullsky_lensing_corrfunc_integrate(lib):
    """Density + RSD + lensing, full sky (test_corrfunc.c:239-280): NONINTEGRATED +
    SINGLE_INTEGRATED + DOUBLE_INTEGRATED of coffe_integrate at five mu against the
    reference's fixtures (5e-5) and the Python API (1e-10)."""
    from coffe_amd.coffe import Coffe
    L, host = lib
    mus = [0.0, 0.2, 0.5, 0.8, 0.95]
    seps = np.array([10, 20, 40, 100, 150]) / H
    cosmo = Coffe(sep=seps, mu=mus, has_density=True, has_rsd=True, has_lensing=True)
    cosmo.set_power_spectrum_linear(*_pk())
    prog = R.Program(host, cosmo, output_type=R.CORRFUNC)
    print("coffe_integrate full-sky lensing corrfunc", flush=True)
    ia = _integrals(L, prog)
    py = {(round(r.mu, 6), round(r.r, 9)): r.value for r in cosmo.compute_corrfunc_bulk()}
    for i, mu in enumerate(mus):
        x, y = np.loadtxt(os.path.join(GOLD, f"benchmark_std_len_corrfunc{i}.dat")).T
        got = []
        for r in seps:
            got.append(sum(L.coffe_integrate(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(ia), 1.0,
                                             0.0, 0.0, float(r), mu, 0, kind, R.CORRFUNC)
                           for kind in (R.NONINTEGRATED, R.SINGLE_INTEGRATED, R.DOUBLE_INTEGRATED)))
            want = py[(round(mu, 6), round(float(r), 9))]
            assert abs(got[-1] - want) <= 1e-10 * abs(want), (mu, r, got[-1], want)
        assert np.allclose(got, y, rtol=5e-5, atol=0), (mu, got, y)
    L.coffe_integrals_free(ctypes.byref(ia))
