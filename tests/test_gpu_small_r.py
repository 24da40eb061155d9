"""GPU parity of the small-separation table nodes (quad_kernels.hip) and of the
reference's full (N + 43)-point tables (coffe_integral_table) against the oracle:
the same QAG-GK61 bisection, so the nodes agree to rounding (1e-10 relative, or
1e-12 of the table's largest small-r node near zero crossings)."""
import os

import numpy as np
import pytest

import oracle as O
from tests.synth import batch_spectra

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
H = 0.67
TABLES = [(0, 0), (0, 2), (0, 4), (1, 1), (1, 3), (2, 0), (2, 2), (3, 1)]


def _close(a, b, rel=1e-10, floor=1e-12):
    scale = np.max(np.abs(b))
    return np.all(np.abs(a - b) <= rel * np.abs(b) + floor * scale)


@pytest.mark.parametrize("nl", TABLES + [(-1, 0), (4, 2), (0, 1)])
def test_integral_table_vs_oracle(engine, nl):
    from coffe_amd.engine import integral_table
    n, l = nl
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    k, pk = k * H, pk / H ** 3
    print(f"integral_table n={n} l={l}", flush=True)
    gx, gy = integral_table(n, l, k, pk, 10000)
    ox, oy = O.integral_table(n, l, k, pk, 10000, small_r=True)
    assert np.array_equal(gx[:43], ox[:43])
    assert np.allclose(gx[43:], ox[43:], rtol=1e-14, atol=0)
    assert _close(gy[:43], oy[:43]), (n, l, np.max(np.abs(gy[:43] / oy[:43] - 1)))
    big = np.abs(oy[43:]) > 1e-6 * np.max(np.abs(oy[43:]))
    assert np.max(np.abs(gy[43:][big] / oy[43:][big] - 1)) < 1e-10


def test_batch_small_r_vs_oracle(engine):
    import torch
    from coffe_amd.engine import Batch
    specs = batch_spectra(5, seed=21, lengths=(1024, 4096, 16384))
    sep = np.linspace(20, 200, 8)
    b = Batch(engine, [len(s[0]) for s in specs], [len(s[0]) for s in specs], [s[0][0] for s in specs],
              [s[0][-1] for s in specs], TABLES, sep, 1, 1)
    k = torch.tensor(np.concatenate([s[0] for s in specs]), device="cuda")
    p = torch.tensor(np.concatenate([s[1] for s in specs]), device="cuda")
    print("batch small r", [len(s[0]) for s in specs], flush=True)
    got = b.small_r(k, p).cpu().numpy()
    b.close()
    for s, (ks, ps, _) in enumerate(specs):
        for t, (n, l) in enumerate(TABLES):
            _, oy = O.integral_table(n, l, ks, ps, 1024, small_r=True)
            assert _close(got[s, t], oy[:43]), (s, n, l)


def test_integrals_renormalizable_dropin(engine):
    """The drop-in of coffe_integrals_renormalizable (integrals.c:459-573) reads the
    spectrum from the reference's coffe_interpolation / gsl_spline layout, mallocs the
    (N + 43)-point table and returns exactly coffe_integral_table's values."""
    import ctypes
    from coffe_amd._lib import load_library
    from coffe_amd.engine import integral_table
    from tests.test_abi import _GslSplineView
    L = load_library()
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    k, pk = np.ascontiguousarray(k * H), np.ascontiguousarray(pk / H ** 3)
    interp, keep = _GslSplineView.make(k, pk)
    f = L.coffe_integrals_renormalizable
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    for n, l in ((0, 0), (1, 3), (2, 0)):
        print(f"coffe_integrals_renormalizable n={n} l={l}", flush=True)
        ox, oy = ctypes.POINTER(ctypes.c_double)(), ctypes.POINTER(ctypes.c_double)()
        n_out = ctypes.c_size_t(0)
        rc = f(ctypes.byref(ox), ctypes.byref(oy), 2048, ctypes.byref(n_out), ctypes.byref(interp), n, l, 0, 0,
               float(k[0]), float(k[-1]))
        assert rc == 0 and n_out.value == 2048 + 43
        x = np.ctypeslib.as_array(ox, shape=(n_out.value,)).copy()
        y = np.ctypeslib.as_array(oy, shape=(n_out.value,)).copy()
        libc.free(ctypes.cast(ox, ctypes.c_void_p))
        libc.free(ctypes.cast(oy, ctypes.c_void_p))
        tx, ty = integral_table(n, l, k, pk, 2048)
        assert np.array_equal(x, tx) and np.array_equal(y, ty)
