"""The integrated terms on the GPU (coffe_integrated_corrfunc, k_integrated_single and
k_integrated_double): the kernels against the oracle (oracle/integrated.py, the same
product rules) for every group of contributions with unequal populations, and the
Coffe API (full-sky lensing, g4, g5, their cross terms with the local ones) against the
reference's fixtures the way its tests hold them (test_coffe.py:235-303, rtol 5e-4)."""
import os

import numpy as np
import pytest

import oracle as O
from oracle import integrated as OI

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
H = 0.67
SEPS = np.array([10, 20, 40, 100, 150]) / H
MUS = [0.0, 0.2, 0.5, 0.8, 0.95]
ALL = ["den", "rsd", "d1", "d2", "g1", "g2", "g3", "len", "g4", "g5"]


def _pk():
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    return k * H, pk / H ** 3


@pytest.fixture(scope="module")
def shared():
    """Oracle tables (N = 4096) and populations whose biases all differ and vary in z."""
    from coffe_amd import background as B
    from coffe_amd.engine import COFFE_SEP
    k, pk = _pk()
    zb = np.linspace(0, 10, 11)
    s1, s2 = 0.1 + 0.03 * zb, 0.25 - 0.01 * zb
    fe1, fe2 = 0.5 * np.ones(11), -0.2 + 0.05 * zb
    b1, b2 = 1.2 + 0.3 * zb, 0.9 + 0.1 * zb
    bg = B.compute(omega_cdm=0.25, omega_baryon=0.05, h=H, bias_z=zb, s1=s1, s2=s2, fevo1=fe1, fevo2=fe2)
    tabs = {key: O.integral_table(key[0], key[1], k, pk, 4096, small_r=True) for key in O.TABLE_ORDER}
    chi = np.arange(201) / 200 * O.akima(bg["z"], bg["chi"], [1.7])[0]
    i40, r0, ren = O.divergent_tables(k, pk, COFFE_SEP, chi)
    tabs[(4, 0)] = (COFFE_SEP, i40)
    x0, y0 = COFFE_SEP.copy(), r0.copy()
    x0[0], y0[0] = 0.0, 0.0
    inp = OI.Inputs(tabs, bg, 0.30, H / 2997.92458, zb, b1, b2, zb, s1, s2, fe1, fe2, div=(COFFE_SEP, i40, r0),
                    ren=(chi, ren))
    bgd = dict(z=bg["z"], chi=bg["chi"], D1=bg["D1"], a=bg["a"], f=bg["f"], cH=bg["conformal_Hz"], G1=bg["G1"],
               G2=bg["G2"])
    pops = (zb, b1, b2, zb, s1, s2, fe1, fe2)
    return inp, tabs, bgd, pops, (chi, ren), (x0, y0)


@pytest.mark.parametrize("flags", [["len"], ["g4"], ["g5"], ["len", "g4", "g5"], ["den", "rsd", "d1", "len"],
                                   ["d2", "g1", "g2", "g3", "len"], ["den", "rsd", "d1", "d2", "g1", "g4", "g5"],
                                   ALL, ALL + ["only_cross"], ["den", "rsd", "len", "g4", "flat_ll", "flat_dl"]])
def test_kernel_vs_oracle(engine, shared, flags):
    """Both integrals at 2 redshifts x 5 mu (both signs, near +-1) x 3 separations, to
    1e-9 of the largest value of each."""
    from coffe_amd.engine import integrated_corrfunc
    inp, tabs, bgd, pops, ren, ren0 = shared
    zs, mus, seps = [0.8, 1.4], [-0.9, -0.3, 0.0, 0.45, 0.97], np.array([12.0, 60.0, 210.0])
    print("integrated kernel", flags, flush=True)
    gs, gd = integrated_corrfunc(bgd, *pops, 0.30, H / 2997.92458, tabs, flags, zs, mus, seps, ren=ren, ren0=ren0)
    fl = {f: 1 for f in flags}
    ws = np.array([[[OI.corrfunc_single(inp, fl, z, m, s) for s in seps] for m in mus] for z in zs])
    wd = np.array([[[OI.corrfunc_double(inp, fl, z, m, s) for s in seps] for m in mus] for z in zs])
    for got, want in ((gs, ws), (gd, wd)):
        scale = np.max(np.abs(want))
        assert np.max(np.abs(got - want)) <= 1e-9 * scale + 1e-300, (flags, np.max(np.abs(got - want)) / scale)


def test_kernel_rejects_bad_inputs(engine, shared):
    """A table the flags need, an unknown flag bit, a mu outside [-1, 1] or along the line
    of sight (|mu| > 1 - 1e-10) is refused."""
    from coffe_amd import EngineError
    from coffe_amd.engine import integrated_corrfunc
    inp, tabs, bgd, pops, ren, ren0 = shared
    args = (bgd, *pops, 0.30, H / 2997.92458)
    no31 = {k: v for k, v in tabs.items() if k != (3, 1)}
    with pytest.raises(EngineError):
        integrated_corrfunc(*args, no31, ["len", "g4"], [1.0], [0.3], [50.0], ren=ren, ren0=ren0)
    with pytest.raises(EngineError):
        integrated_corrfunc(*args, tabs, ["g4"], [1.0], [0.3], [50.0])       # no renormalisation grid
    with pytest.raises(EngineError):
        integrated_corrfunc(*args, tabs, ["len"], [1.0], [1.5], [50.0])
    with pytest.raises(EngineError) as e:          # along the line of sight: COFFE_EDOMAIN
        integrated_corrfunc(*args, tabs, ["len"], [1.0], [1.0], [50.0])
    assert e.value.code == 5
    s_, d_ = integrated_corrfunc(*args, tabs, ["len"], [1.0], [1.0 - 2e-10], [50.0])
    assert np.isfinite(d_).all()
    with pytest.raises(ValueError):
        integrated_corrfunc(*args, tabs, ["len", "g6"], [1.0], [0.3], [50.0])


def _coffe(**kw):
    from coffe_amd.coffe import Coffe
    c = Coffe(**kw)
    c.set_power_spectrum_linear(*_pk())
    return c


@pytest.mark.parametrize("pre,flag", [("len", "has_lensing"), ("g4", "has_g4"), ("g5", "has_g5")])
def test_corrfunc_vs_reference(engine, pre, flag):
    """test_coffe.py:235-268 for lensing, g4 and g5 (the reference's test sets each alone)."""
    c = _coffe(sep=SEPS, mu=MUS, has_density=False)
    setattr(c, flag, True)
    print("Coffe corrfunc", pre, flush=True)
    result = c.compute_corrfunc_bulk()
    for i, mu in enumerate(MUS):
        x, y = np.loadtxt(os.path.join(GOLD, f"benchmark_{pre}_corrfunc{i}.dat")).T
        rows = [r for r in result if r.mu == mu]
        assert np.allclose([r.r * H for r in rows], x)
        assert np.allclose([r.value for r in rows], y, rtol=5e-4, atol=0), (pre, mu)


def test_density_rsd_lensing_vs_reference(engine):
    """Density + RSD + lensing (test_corrfunc.c:239-280 and test_multipoles.c:271-330):
    the correlation function within 5e-5; the multipoles within the fixtures' own 3D
    integration error at 10-40 Mpc/h (the lensing-lensing part, see the oracle's tests)
    and 5e-4 beyond."""
    c = _coffe(sep=SEPS, mu=MUS, l=[0, 2, 4], has_density=True, has_rsd=True, has_lensing=True)
    print("Coffe std + lensing", flush=True)
    cf = c.compute_corrfunc_bulk()
    for i, mu in enumerate(MUS):
        x, y = np.loadtxt(os.path.join(GOLD, f"benchmark_std_len_corrfunc{i}.dat")).T
        assert np.allclose([r.value for r in cf if r.mu == mu], y, rtol=5e-5, atol=0), mu
    mp = c.compute_multipoles_bulk()
    for l in (0, 2, 4):
        x, y = np.loadtxt(os.path.join(GOLD, f"benchmark_std_len_multipoles{l}.dat")).T
        got = np.array([r.value for r in mp if r.l == l])
        assert np.allclose(got[3:], y[3:], rtol=5e-4, atol=0), (l, got / y - 1)
        assert np.allclose(got, y, rtol=2e-2, atol=0), (l, got / y - 1)


def test_lensing_multipoles(engine):
    """Full-sky lensing multipoles: the flat-sky lensing fixtures at 20 and 40 Mpc/h
    within 1e-3, the full-sky monopole fixtures at 100 and 150 Mpc/h within 5e-4."""
    c = _coffe(sep=SEPS, l=[0, 2, 4], has_density=False, has_lensing=True)
    print("Coffe lensing multipoles", flush=True)
    mp = c.compute_multipoles_bulk()
    for l in (0, 2, 4):
        got = np.array([r.value for r in mp if r.l == l])
        xf, yf = np.loadtxt(os.path.join(GOLD, f"benchmark_flatsky_lensing_lensing_multipoles{l}.dat")).T
        for sep_h in (20, 40):
            v = got[int(np.argmin(np.abs(SEPS * H - sep_h)))]
            ref = yf[int(np.argmin(np.abs(xf - sep_h)))]
            assert abs(v - ref) <= 1e-3 * abs(ref), (l, sep_h, v / ref - 1)
        if l == 0:
            x, y = np.loadtxt(os.path.join(GOLD, "benchmark_len_multipoles0.dat")).T
            assert np.allclose(got[3:], y[3:], rtol=5e-4, atol=0)


@pytest.mark.parametrize("pre,flag", [("g4", "has_g4"), ("g5", "has_g5")])
def test_g4_g5_multipoles_vs_reference(engine, pre, flag):
    """test_coffe.py:270-303 for g4 and g5: the monopoles within 5e-4; the quadrupole and
    hexadecapole, 1e-4 .. 1e-7 of the monopole, within 5e-4 of the monopole (the
    reference's np.allclose with its default atol 1e-8 passes them all)."""
    c = _coffe(sep=SEPS, l=[0, 2, 4], has_density=False)
    setattr(c, flag, True)
    print("Coffe multipoles", pre, flush=True)
    mp = c.compute_multipoles_bulk()
    mono = np.loadtxt(os.path.join(GOLD, f"benchmark_{pre}_multipoles0.dat"))[:, 1]
    for l in (0, 2, 4):
        x, y = np.loadtxt(os.path.join(GOLD, f"benchmark_{pre}_multipoles{l}.dat")).T
        got = np.array([r.value for r in mp if r.l == l])
        assert np.all(np.abs(got - y) <= 5e-4 * np.abs(mono)), (pre, l, (got - y) / mono)
        assert np.allclose(got, y, rtol=5e-4)


def test_flat_switches_keep_full_sky_potential_terms(engine):
    """With has_flatsky_nonlocal and has_flatsky_local_nonlocal the lensing-lensing and
    density-lensing terms take their flat-sky forms while g4 and its cross terms stay
    full sky: the sum equals the flat-sky lensing result plus the integrated kernel's
    g4 terms computed with the flat switches (lensing-lensing and density-lensing left
    out)."""
    kw = dict(sep=SEPS[1:], mu=MUS, has_density=True, has_lensing=True, has_flatsky_nonlocal=True,
              has_flatsky_local_nonlocal=True)
    both = _coffe(**kw, has_g4=True)
    lens_only = _coffe(**kw)
    print("flat switches with g4", flush=True)
    a = np.array([r.value for r in both.compute_corrfunc_bulk()])
    b = np.array([r.value for r in lens_only.compute_corrfunc_bulk()])
    plan = both._lensing_plan()
    assert plan["flat_ll"] and plan["flat_dl"] and {"g4", "flat_ll", "flat_dl"} <= plan["flags"]
    extra = both._integrated_corrfunc(both.z_mean, both.sep, both.mu).reshape(-1)
    assert np.allclose(a, b + extra, rtol=1e-12, atol=1e-20)


def test_average_multipoles_with_integrated_terms(engine):
    """Redshift-averaged multipoles with lensing and g4 (signal.c:312-470, 674-682,
    759-767): the local part equals the averaged multipoles without them; the integrated
    part of a bin of width 0.1 equals h H0 / H(z) times the multipoles at its centre
    within 3e-3 (the bin's curvature: a sharp z weighting reduces to the value there);
    the reference ABI (coffe_average_multipoles_init) equals the Python API to 1e-10."""
    import ctypes
    from coffe_amd.coffe import COFFE_H0
    from tests import refabi_host as R
    seps = np.array([30.0, 100.0])
    kw = dict(sep=seps, l=[0, 2], z_min=[0.95], z_max=[1.05], z_mean=[1.0], has_density=True)
    full = _coffe(**kw, has_lensing=True, has_g4=True)
    local = _coffe(**kw)
    print("averaged multipoles with lensing and g4", flush=True)
    a = np.array([r.value for r in full.compute_average_multipoles_bulk()])
    b = np.array([r.value for r in local.compute_average_multipoles_bulk()])
    part = (a - b).reshape(2, 2)                                  # [l][r]
    mp = (np.array([r.value for r in full.compute_multipoles_bulk()]) -
          np.array([r.value for r in local.compute_multipoles_bulk()])).reshape(2, 2)
    w = H * COFFE_H0 / full.hubble_rate(1.0)
    assert np.allclose(part, w * mp, rtol=3e-3, atol=0), part / (w * mp) - 1
    # the reference ABI
    from coffe_amd._lib import load_library
    L = R.declare(load_library())
    host = R.Host(L)
    try:
        prog = R.Program(host, full, output_type=R.AVERAGE_MULTIPOLES)
        ia = R.IntegralArray()
        assert L.coffe_integrals_init(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(ia)) == 0
        ram = R.AverageMultipolesArray()
        assert L.coffe_average_multipoles_init(ctypes.byref(prog.par), ctypes.byref(prog.bg), ctypes.byref(ia),
                                               ctypes.byref(ram)) == 0
        got = np.array([ram.array[i].value for i in range(ram.size)])
        assert np.allclose(got, a, rtol=1e-10, atol=0), got / a - 1
        L.coffe_average_multipoles_free(ctypes.byref(ram))
        L.coffe_integrals_free(ctypes.byref(ia))
    finally:
        host.close()
