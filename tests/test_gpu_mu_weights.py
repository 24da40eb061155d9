"""GPU parity of the mu-quadrature weights (coffe_mu_weights) against the oracle's
multipoles (functions.c:89-1484, signal.c:83-111, 521-542, multipoles.c:48-136):

- full-sky weights against the reference's formulas evaluated in binary128
  (orc_mu_weights_hp) under the same rule: the engine uses their cancellation-free
  forms (fullsky_forms.h), so the (chi/r)^4 terms that cost an FP64 evaluation of the
  reference's own form up to ~1e-5 at 5 Mpc/h no longer limit the agreement;
- flat-sky weights against the closed form of the den+RSD multipoles, 1e-12;
- the whole GPU path (tables + weights + projection) against the reference's own
  multipole fixtures (test_coffe.py:270-303), rtol 5e-4 as there.
"""
import os

import numpy as np
import pytest

import oracle as O
from coffe_amd import background as B

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
H = 0.67
TABLES = [(0, 0), (0, 2), (0, 4), (1, 1), (1, 3), (2, 0), (2, 2), (3, 1)]
FLAGS = {"den": dict(den=1, rsd=0, d1=0), "rsd": dict(den=0, rsd=1, d1=0), "d1": dict(den=0, rsd=0, d1=1),
         "std": dict(den=1, rsd=1, d1=0), "std_d1": dict(den=1, rsd=1, d1=1)}


@pytest.fixture(scope="module")
def setup():
    k, pk = np.loadtxt(os.path.join(GOLD, "PkL_CLASS.dat")).T
    k, pk = k * H, pk / H ** 3
    bg = B.compute(omega_cdm=0.25, omega_baryon=0.05, h=H)
    bgo = dict(z=bg["z"], chi=bg["chi"], D1=bg["D1"], f=bg["f"], cH=bg["conformal_Hz"], G1=bg["G1"], G2=bg["G2"])
    tabs = {key: O.integral_table(key[0], key[1], k, pk, 10000, small_r=False) for key in O.TABLE_ORDER}
    return k, pk, bgo, tabs


def _project(A, tabs, sep, scale=False):
    I = np.array([O.akima(tabs[t][0], tabs[t][1], sep) for t in TABLES[:A.shape[2]]])
    if scale:
        return np.einsum("zltr,tr->zlr", np.abs(A), np.abs(I))
    return np.einsum("zltr,tr->zlr", A, I)


@pytest.mark.parametrize("flags", ["std_d1", "den", "rsd", "d1"])
def test_fullsky_weights_vs_oracle_hp(engine, setup, flags):
    """Full-sky weights against the reference's formulas evaluated in binary128
    (orc_mu_weights_hp, pinned to a 50-digit mpmath evaluation) under the same 256-node
    rule, at 5-300 Mpc/h and z = 0.5, 1, 2: every weight to 1e-12 of the largest |weight|
    at its (z, r), and the multipoles (weights times the oracle's tables) to 1e-10
    relative per point plus 1e-12 of the monopole there."""
    from coffe_amd.engine import mu_weights
    k, pk, bgo, tabs = setup
    sep = np.array([5, 7.5, 10, 20, 40, 60, 100, 150, 200, 300]) / H
    zb = np.linspace(0, 10, 11)
    b1, b2 = 1.2 + 0.3 * zb, 1.5 + 0.1 * zb
    zm = [0.5, 1.0, 2.0]
    print(f"mu weights full sky {flags}", flush=True)
    A = mu_weights(engine, bgo, zb, b1, b2, zm, [0, 2, 4], TABLES, sep, flatsky=False, nmu=256,
                   **FLAGS[flags]).cpu().numpy()
    hp = O.mu_weights_hp(bgo, zb, b1, b2, zm, [0, 2, 4], sep, flatsky=False, nmu=256, **FLAGS[flags])
    for iz in range(len(zm)):
        for ir in range(len(sep)):
            scale = np.max(np.abs(hp[iz, :, :, ir]))
            err = np.max(np.abs(A[iz, :, :, ir] - hp[iz, :, :, ir]))
            assert err <= 1e-12 * scale, (flags, zm[iz], sep[ir], err / scale)
    got = _project(A, tabs, sep)
    ref = _project(hp, tabs, sep)
    # density or Doppler alone have l >= 2 only through the wide-angle dependence of the
    # end points' bias, growth and D1 (O((s/chi)^2) of the monopole), where one ulp of
    # those FP64 inputs moves the result by ~1e-8 of itself: the bound adds 1e-12 of the
    # monopole at the same point
    for il, l in enumerate([0, 2, 4]):
        if l == 4 and flags in ("den", "d1"):
            continue   # vanishes in theory
        err = np.abs(got[:, il] - ref[:, il])
        bound = 1e-10 * np.abs(ref[:, il]) + 1e-12 * np.abs(ref[:, 0])
        print(f"  l={l}: max rel {np.max(err / np.abs(ref[:, il])):.2e}, max err/bound {np.max(err / bound):.2e}",
              flush=True)
        assert np.all(err <= bound), (flags, l, np.max(err / bound))


def test_flatsky_weights_closed_form(engine, setup):
    from coffe_amd.engine import mu_weights
    k, pk, bgo, tabs = setup
    sep = np.linspace(5, 300, 24) / H
    zb = np.linspace(0, 10, 11)
    b = np.full(11, 1.5)
    A = mu_weights(engine, bgo, zb, b, b, [1.0], [0, 2, 4], TABLES[:3], sep, den=1, rsd=1, d1=0, flatsky=True,
                   nmu=64).cpu().numpy()
    f = O.akima(bgo["z"], bgo["f"], [1.0])[0]
    D = O.akima(bgo["z"], bgo["D1"], [1.0])[0]
    closed = {(0, 0): D * D * (1.5 ** 2 + 2 * 1.5 * f / 3 + f * f / 5), (2, 1): -D * D * (4 * 1.5 * f / 3 + 4 * f * f / 7),
              (4, 2): D * D * 8 * f * f / 35}
    for il, l in enumerate([0, 2, 4]):
        for t in range(3):
            want = closed.get((l, t), 0.0)
            assert np.allclose(A[0, il, t], want, rtol=1e-12, atol=1e-14 * abs(D * D)), (l, t)


def test_flatsky_weights_vs_oracle_d1(engine, setup):
    from coffe_amd.engine import mu_weights
    k, pk, bgo, tabs = setup
    sep = np.array([10, 40, 100, 200]) / H
    zb = np.linspace(0, 10, 11)
    b1, b2 = 1.2 + 0.3 * zb, 1.5 + 0.1 * zb
    # b1 != b2 at z = 1 gives a dipole; with G1 = G2 the flat-sky octupole vanishes
    A = mu_weights(engine, bgo, zb, b1, b2, [1.0], [0, 1, 2, 4], TABLES, sep, den=1, rsd=1, d1=1, flatsky=True,
                   nmu=256).cpu().numpy()
    got = _project(A, tabs, sep)
    ref = O.multipoles(tabs, bgo, zb, b1, b2, [1.0], [0, 1, 2, 4], sep, den=1, rsd=1, d1=1, flatsky=True, quad=256)
    assert np.min(np.abs(ref)) > 1e-8
    assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-10


@pytest.mark.parametrize("pre", ["den", "rsd", "d1", "std", "std_d1"])
def test_gpu_path_vs_reference_fixtures(engine, setup, pre):
    import torch
    from coffe_amd.engine import Batch, mu_weights
    k, pk, bgo, tabs = setup
    sep = np.array([10, 20, 40, 100, 150]) / H
    zb, b = np.arange(11.0), np.ones(11)
    A = mu_weights(engine, bgo, zb, b, b, [1.0], [0, 2, 4], TABLES, sep, flatsky=False, nmu=256, **FLAGS[pre])
    batch = Batch(engine, [len(k)], [10000], [k[0]], [k[-1]], TABLES, sep, 1, 3)
    print(f"gpu path vs fixtures {pre}", flush=True)
    _, out = batch.forward(torch.tensor(k, device="cuda"), torch.tensor(pk, device="cuda"), A)
    out = out.cpu().numpy()[0, 0]
    batch.close()
    for il, l in enumerate([0, 2, 4]):
        x, y = np.loadtxt(os.path.join(GOLD, f"benchmark_{pre}_multipoles{l}.dat")).T
        assert np.allclose(sep * H, x)
        assert np.allclose(out[il], y, rtol=5e-4), (pre, l)


def test_domain_errors(engine, setup):
    from coffe_amd.engine import mu_weights
    k, pk, bgo, tabs = setup
    zb, b = np.arange(11.0), np.ones(11)
    with pytest.raises(RuntimeError):   # chi(0.01) - sep/2 < 0
        mu_weights(engine, bgo, zb, b, b, [0.01], [0], TABLES, [500.0], flatsky=False)
    with pytest.raises(RuntimeError):   # tables required by RSD missing
        mu_weights(engine, bgo, zb, b, b, [1.0], [0], [(0, 0)], [100.0], den=1, rsd=1, flatsky=False)


def test_async_calls_back_to_back(engine, setup):
    """coffe_mu_weights is asynchronous on its stream: several calls queued back to back,
    each with host inputs that are overwritten right after the call returns (the pinned
    staging takes them at once), must equal the same calls made one at a time."""
    import torch
    from coffe_amd.engine import mu_weights
    _, _, bgo, _ = setup
    sep = np.linspace(20, 200, 16) / H
    zb = np.linspace(0, 10, 11)
    cases = [(1.0 + 0.1 * i, 1.2 + 0.05 * i, [0.6 + 0.2 * i, 1.4]) for i in range(5)]
    ref = []
    for b0, b1, zm in cases:
        A = mu_weights(engine, bgo, zb, np.full(11, b0), np.full(11, b1), zm, [0, 2], TABLES, sep, den=1, rsd=1,
                       d1=1, flatsky=False, nmu=64)
        torch.cuda.synchronize()
        ref.append(A.cpu().numpy())
    outs = []
    bias1, bias2, zmean = np.empty(11), np.empty(11), np.empty(2)
    print("mu weights queued back to back", flush=True)
    for b0, b1, zm in cases:
        bias1[:] = b0
        bias2[:] = b1
        zmean[:] = zm
        outs.append(mu_weights(engine, bgo, zb, bias1, bias2, zmean, [0, 2], TABLES, sep, den=1, rsd=1, d1=1,
                               flatsky=False, nmu=64))
        bias1[:] = np.nan   # the call has taken its inputs
        bias2[:] = np.nan
        zmean[:] = np.nan
    torch.cuda.synchronize()
    for got, want in zip(outs, ref):
        assert np.array_equal(got.cpu().numpy(), want)
