"""Error behaviour of the drop-in entry points and the clean-up paths of every allocating
call, on the GPU.

The reference exit()s on bad input (twofast.c:62, 81-88, 122, 210; twobessel.c:20-23);
the engine's void drop-ins twofast_1bessel / two_sph_bessel record a code that
coffe_engine_last_error() returns instead, and the next good call clears it.  With
coffe_engine_inject_alloc_failure(n) the n-th device allocation fails: batch, covariance
plan and drop-in calls must then return COFFE_ENOMEM, leave nothing behind, and the same
calls must succeed with the injection off."""
import ctypes

import numpy as np
import pytest

from tests.synth import spectrum

pytestmark = pytest.mark.gpu

DP = ctypes.POINTER(ctypes.c_double)
EINVAL, ENOMEM, EUNSUPPORTED, EDOMAIN, ENONFINITE = 1, 2, 4, 5, 6


def _void_twofast(L, k, p, N, kmin=None, kmax=None, l=0.0, nu=0.0):
    k = np.ascontiguousarray(k, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64)
    ox, oy = np.zeros(max(N, 1)), np.zeros(max(N, 1))
    kmin = float(k[0]) if kmin is None else kmin
    kmax = float(k[-1]) if kmax is None else kmax
    L.twofast_1bessel(ox.ctypes.data_as(DP), oy.ctypes.data_as(DP), N, k.ctypes.data_as(DP), p.ctypes.data_as(DP),
                      len(k), l, nu, 1.0, kmin, kmin, kmax, 0)
    return L.coffe_engine_last_error(), oy


def test_twofast_void_error_codes(engine):
    from coffe_amd._lib import load_library
    L = load_library()
    k, p = spectrum(1024)
    assert _void_twofast(L, k, p, 1024)[0] == 0
    print("void twofast: odd N", flush=True)
    assert _void_twofast(L, k, p, 1023)[0] == EUNSUPPORTED                 # the reference's FFT needs even N
    assert _void_twofast(L, k, p, 1 << 22)[0] == EUNSUPPORTED              # above the general path's 2^21
    bad = p.copy()
    bad[17] = np.nan
    assert _void_twofast(L, k, bad, 1024)[0] == ENONFINITE                 # twofast.c:122 exits
    badk = k.copy()
    badk[3] = np.inf
    assert _void_twofast(L, badk, p, 1024)[0] == ENONFINITE
    print("void twofast: kmin >= kmax, inverted window", flush=True)
    assert _void_twofast(L, k, p, 1024, kmin=float(k[-1]), kmax=float(k[0]))[0] == EINVAL
    assert _void_twofast(L, k, p, 1024, kmin=float(k[10]), kmax=float(k[10]))[0] == EINVAL
    # a range shorter than the window's e^{0.46} edges: xl > xr (twofast.c:81-88 exits)
    assert _void_twofast(L, k, p, 1024, kmin=float(k[0]), kmax=float(k[0]) * 2.0)[0] == EDOMAIN
    assert _void_twofast(L, k[:4], p[:4], 1024)[0] == EINVAL              # fewer than 5 samples
    st, oy = _void_twofast(L, k, p, 1024)
    assert st == 0 and np.all(np.isfinite(oy))                             # a good call clears the error


def test_two_sph_bessel_void_error_codes(engine):
    from coffe_amd._lib import load_library
    from coffe_amd.covariance import Config
    L = load_library()
    L.two_sph_bessel.argtypes = [DP, DP, ctypes.POINTER(DP), ctypes.c_long, ctypes.c_long, ctypes.c_void_p, DP, DP,
                                 ctypes.POINTER(DP)]
    L.two_sph_bessel.restype = None

    def call(N1, N2, diag=True):
        k1 = 1e-3 * 1e3 ** (np.arange(N1) / N1)
        k2 = 1e-3 * 1e3 ** (np.arange(N2) / N2)
        fk = np.zeros((N1, N2))
        if diag:
            for i in range(min(N1, N2)):
                fk[i, i] = k1[i] ** 3
        else:
            fk[:] = 1.0
        r1, r2, res = np.empty(N1), np.empty(N2), np.empty((N1, N2))
        rin = (DP * N1)(*[fk[i].ctypes.data_as(DP) for i in range(N1)])
        rout = (DP * N1)(*[res[i].ctypes.data_as(DP) for i in range(N1)])
        cfg = Config(0.0, 2.0, 1.01, 1.01, 0.25, N1)
        L.two_sph_bessel(k1.ctypes.data_as(DP), k2.ctypes.data_as(DP), rin, N1, N2, ctypes.byref(cfg),
                         r1.ctypes.data_as(DP), r2.ctypes.data_as(DP), rout)
        return L.coffe_engine_last_error(), res

    assert call(64, 64)[0] == 0
    print("void two_sph_bessel: odd, unequal, dense", flush=True)
    assert call(63, 63)[0] == EINVAL             # twobessel.c exits on odd lengths
    assert call(64, 32)[0] == EUNSUPPORTED       # only the square grid of covariance.c
    assert call(64, 64, diag=False)[0] == EUNSUPPORTED
    st, res = call(64, 64)
    assert st == 0 and np.all(np.isfinite(res))


def test_allocation_failures_clean_up(engine):
    """Every allocation of a batch, a covariance plan and the drop-in's workspace made to
    fail in turn: each call returns ENOMEM, and afterwards the same calls succeed and give
    the same results as before."""
    import torch
    from coffe_amd._lib import EngineError, load_library
    from coffe_amd.covariance import CovPlan
    from coffe_amd.engine import Batch, Engine, twofast_1bessel
    L = load_library()
    k_np, p_np = spectrum(2048)
    sep = np.linspace(10, 200, 32) / 0.67
    tables = [(0, 0), (0, 2), (0, 4)]
    k = torch.tensor(k_np, device="cuda")
    p = torch.tensor(p_np, device="cuda")
    # a fresh engine: its grid cache is empty, so the grid tables are allocated too
    eng = Engine(0)
    ref_b = None
    fails = 0
    try:
        for n in range(1, 60):
            L.coffe_engine_inject_alloc_failure(n)
            try:
                b = Batch(eng, [2048], [2048], [k_np[0]], [k_np[-1]], tables, sep, 1, 3)
            except EngineError as err:
                assert err.code == ENOMEM, err
                fails += 1
                continue
            finally:
                L.coffe_engine_inject_alloc_failure(0)
            itab, _ = b.forward(k, p)
            ref_b = itab.cpu().numpy()
            b.close()
            break
        print(f"batch: {fails} injected failures before success", flush=True)
        assert fails >= 10 and ref_b is not None
        b = Batch(eng, [2048], [2048], [k_np[0]], [k_np[-1]], tables, sep, 1, 3)
        assert np.array_equal(b.forward(k, p)[0].cpu().numpy(), ref_b)
        b.close()
        fails = 0
        for n in range(1, 60):
            L.coffe_engine_inject_alloc_failure(n)
            try:
                plan = CovPlan(eng, 256, k_np[0], k_np[-1], [0, 2], sep)
            except EngineError as err:
                assert err.code == ENOMEM, err
                fails += 1
                continue
            finally:
                L.coffe_engine_inject_alloc_failure(0)
            got = plan.integrals(k, p, [2048]).cpu().numpy()
            plan.close()
            break
        print(f"covariance plan: {fails} injected failures before success", flush=True)
        assert fails >= 5
        plan = CovPlan(eng, 256, k_np[0], k_np[-1], [0, 2], sep)
        assert np.array_equal(plan.integrals(k, p, [2048]).cpu().numpy(), got)
        plan.close()
    finally:
        L.coffe_engine_inject_alloc_failure(0)
        eng.close()
    # the drop-in (default engine): a failure of its first allocation, then a good call
    L.coffe_engine_inject_alloc_failure(1)
    ox, oy = np.zeros(3750), np.zeros(3750)
    st = L.twofast_1bessel_status(ox.ctypes.data_as(DP), oy.ctypes.data_as(DP), 3750,
                                  k_np.ctypes.data_as(DP), p_np.ctypes.data_as(DP), len(k_np), 1.0, 1.0, 1.0,
                                  float(k_np[0]), float(k_np[0]), float(k_np[-1]))
    L.coffe_engine_inject_alloc_failure(0)
    assert st == ENOMEM
    rx, ry = twofast_1bessel(k_np, p_np, 3750, 1, 1)
    assert np.all(np.isfinite(ry))
