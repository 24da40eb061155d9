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


@pytest.mark.parametrize("N", [1024, 4096, 10000, 16384])
@pytest.mark.parametrize("nl", [(0, 0), (0, 2), (2, 0), (3, 1), (0.5, -0.5)])
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
