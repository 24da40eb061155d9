"""GPU parity of the FFTLog engine against the oracle (CPU restatement of
twofast.c / integrals.c).  Tolerance: 1e-10 relative on I_l^n(r) where the value
exceeds 1e-6 of the table's maximum."""
import numpy as np
import pytest

import oracle as O
from tests.synth import batch_spectra, spectrum

pytestmark = pytest.mark.gpu

TABLES = [(0, 0), (0, 2), (0, 4), (1, 1), (1, 3), (2, 0), (2, 2), (3, 1)]


def _rel(a, b, floor_frac=1e-6):
    scale = np.max(np.abs(b))
    m = np.abs(b) > floor_frac * scale
    return np.max(np.abs(a[m] - b[m]) / np.abs(b[m])) if m.any() else 0.0


# 6144 = 2^11 * 3 and 10000 = 2^4 * 5^4 take the radix-2, -3 and -5 passes of the transforms
@pytest.mark.parametrize("N", [1024, 4096, 6144, 10000, 16384])
@pytest.mark.parametrize("nl", [(0, 0), (0, 2), (0, 4), (2, 0), (3, 1), (0.5, -0.5)])
def test_twofast_dropin(engine, N, nl):
    from coffe_amd.engine import twofast_1bessel
    nu, l = nl
    if l < 0:
        l, nu = 0.5, 0.5
    k, p = spectrum(4096)
    print(f"twofast N={N} l={l} nu={nu}", flush=True)
    rx, ry = twofast_1bessel(k, p, N, l, nu)
    ox, oy = O.twofast_1bessel(k, p, N, l, nu)
    assert np.array_equal(rx, ox)
    assert _rel(ry, oy) < 1e-10


def test_batch_tables_full(engine):
    import torch
    from coffe_amd.engine import Batch
    specs = batch_spectra(6, seed=3)
    sep = np.linspace(5, 300, 64) / 0.67
    b = Batch(engine, [len(s[0]) for s in specs], [len(s[0]) for s in specs], [s[0][0] for s in specs],
              [s[0][-1] for s in specs], TABLES, sep, 1, 3)
    k = torch.tensor(np.concatenate([s[0] for s in specs]), device="cuda")
    p = torch.tensor(np.concatenate([s[1] for s in specs]), device="cuda")
    print("batch full tables", [len(s[0]) for s in specs], flush=True)
    full = b.tables_full(k, p).cpu().numpy()
    itab, _ = b.forward(k, p)
    itab = itab.cpu().numpy()
    for s, (ks, ps, meta) in enumerate(specs):
        N = len(ks)
        for t, (n, l) in enumerate(TABLES):
            tx, ty = O.integral_table(n, l, ks, ps, N, small_r=False)
            assert _rel(full[s, t, :N], ty[43:]) < 1e-10, (s, n, l)
            ref = O.akima(tx, ty, sep)
            assert _rel(itab[s, t], ref, 0) < 1e-10, (s, n, l)


@pytest.mark.parametrize("n_in", [5, 9, 100, 513, 514, 1100])
def test_small_and_chunked_inputs(engine, n_in):
    """The spline of the input runs in chunks of 512 intervals with 64-node halos, one
    wavefront each (11 rows of the system per lane): inputs of a few nodes (most lanes
    idle), one chunk exactly, and several chunks must all reproduce the oracle, in the
    drop-in call and in a batch (with its adjoint's ranges)."""
    import torch
    from coffe_amd.engine import Batch, twofast_1bessel
    k, p = spectrum(n_in)
    print(f"twofast n_in={n_in}", flush=True)
    rx, ry = twofast_1bessel(k, p, 1024, 0, 0)
    ox, oy = O.twofast_1bessel(k, p, 1024, 0, 0)
    assert np.array_equal(rx, ox)
    assert _rel(ry, oy) < 1e-10
    sep = np.linspace(5, 300, 32) / 0.67
    b = Batch(engine, [n_in], [1024], [k[0]], [k[-1]], [(0, 0), (0, 2)], sep, 1, 2)
    kt, pt = torch.tensor(k, device="cuda"), torch.tensor(p, device="cuda")
    full = b.tables_full(kt, pt).cpu().numpy()
    for t, (n, l) in enumerate([(0, 0), (0, 2)]):
        tx, ty = O.integral_table(n, l, k, p, 1024, small_r=False)
        assert _rel(full[0, t, :1024], ty[43:]) < 1e-10, (n, l)
    # the adjoint: the map P -> multipoles is positively homogeneous of degree 1, so
    # <gpk, P> = <gout, out>
    A = torch.tensor(np.random.default_rng(4).standard_normal((1, 2, 2, len(sep))), device="cuda")
    gout = torch.tensor(np.random.default_rng(6).standard_normal((1, 1, 2, len(sep))), device="cuda")
    _, out, gpk = b.vjp(kt, pt, A, gout)
    lhs = float(torch.dot(gpk, pt))
    rhs = float(torch.sum(gout * out))
    assert abs(lhs - rhs) <= 1e-10 * (abs(rhs) + float(torch.sum(torch.abs(gout * out))))
