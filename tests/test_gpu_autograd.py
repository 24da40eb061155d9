"""Gradients of the differentiable multipoles (coffe_amd.autograd):
- d/dP through autograd equals the engine's adjoint (same kernels, bit-identical);
- the (b, f) polynomial of the weight groups equals weights computed directly with
  constant bias and growth-rate tables (1e-12), full and flat sky;
- d/db, d/df against central differences, exact for a quadratic up to rounding (1e-8);
- sigma8: dL/dsigma8 = 2 L / sigma8 at fixed P, and the sigma8-normalised multipoles
  are homogeneous of degree 0 in P, so <dL/dP, P> = 0 (Euler)."""
import numpy as np
import pytest

from tests.synth import batch_spectra

pytestmark = pytest.mark.gpu

H = 0.67
TABLES = [(0, 0), (0, 2), (0, 4), (1, 1), (1, 3), (2, 0), (2, 2), (3, 1)]


@pytest.fixture(scope="module")
def bg():
    from coffe_amd import background as B
    b = B.compute(omega_cdm=0.26, omega_baryon=0.05, h=H)
    return dict(z=b["z"], chi=b["chi"], D1=b["D1"], f=b["f"], cH=b["conformal_Hz"], G1=b["G1"], G2=b["G2"])


@pytest.fixture(scope="module")
def layout(engine):
    import torch
    from coffe_amd.engine import Batch
    specs = batch_spectra(4, seed=11, lengths=(1024, 2048, 4096))
    sep = np.linspace(5, 300, 16) / H
    b = Batch(engine, [len(s[0]) for s in specs], [len(s[0]) for s in specs], [s[0][0] for s in specs],
              [s[0][-1] for s in specs], TABLES, sep, 2, 3)
    k = torch.tensor(np.concatenate([s[0] for s in specs]), device="cuda")
    p = torch.tensor(np.concatenate([s[1] for s in specs]), device="cuda")
    yield b, k, p, sep
    b.close()


@pytest.mark.parametrize("flatsky", [False, True])
def test_bf_polynomial_equals_direct_weights(engine, bg, layout, flatsky):
    from coffe_amd.autograd import LocalWeights
    from coffe_amd.engine import mu_weights
    b, k, p, sep = layout
    lw = LocalWeights(engine, bg, [0.8, 1.6], [0, 2, 4], TABLES, sep, den=1, rsd=1, d1=1, flatsky=flatsky,
                      f_override=True)
    A = lw.weights(b=1.3, f=0.7).cpu().numpy()
    bgf = dict(bg)
    bgf["f"] = np.full_like(bg["f"], 0.7)
    zb = np.linspace(0, bg["z"][-1], 11)
    direct = mu_weights(engine, bgf, zb, np.full(11, 1.3), np.full(11, 1.3), [0.8, 1.6], [0, 2, 4], TABLES, sep,
                        den=1, rsd=1, d1=1, flatsky=flatsky).cpu().numpy()
    assert np.max(np.abs(A - direct)) <= 1e-12 * np.max(np.abs(direct))


def test_pk_gradient_is_engine_adjoint(engine, bg, layout):
    import torch
    from coffe_amd.autograd import LocalWeights, multipoles
    b, k, p, sep = layout
    lw = LocalWeights(engine, bg, [0.8, 1.6], [0, 2, 4], TABLES, sep, den=1, rsd=1, d1=1)
    A = lw.weights(b=1.5)
    gout = torch.randn((4, 2, 3, len(sep)), dtype=torch.float64, device="cuda", generator=torch.Generator(
        device="cuda").manual_seed(5))
    pk = p.clone().requires_grad_(True)
    M = multipoles(b, k, pk, A)
    (M * gout).sum().backward()
    _, out, gpk = b.vjp(k, p, A, gout)
    assert torch.equal(M.detach(), out)
    assert torch.equal(pk.grad, gpk)


@pytest.mark.parametrize("flatsky", [False, True])
def test_bf_gradients_central_difference(engine, bg, layout, flatsky):
    import torch
    from coffe_amd.autograd import LocalWeights, multipoles
    b, k, p, sep = layout
    lw = LocalWeights(engine, bg, [0.8, 1.6], [0, 2, 4], TABLES, sep, den=1, rsd=1, d1=1, flatsky=flatsky,
                      f_override=True)
    gout = torch.randn((4, 2, 3, len(sep)), dtype=torch.float64, device="cuda",
                       generator=torch.Generator(device="cuda").manual_seed(7))

    def loss(bv, fv):
        return (multipoles(b, k, p, lw.weights(b=bv, f=fv)) * gout).sum()

    bt = torch.tensor(1.4, dtype=torch.float64, device="cuda", requires_grad=True)
    ft = torch.tensor(0.6, dtype=torch.float64, device="cuda", requires_grad=True)
    L = loss(bt, ft)
    L.backward()
    h = 1e-3
    with torch.no_grad():
        db = (loss(1.4 + h, 0.6) - loss(1.4 - h, 0.6)) / (2 * h)
        df = (loss(1.4, 0.6 + h) - loss(1.4, 0.6 - h)) / (2 * h)
    assert abs(bt.grad.item() - db.item()) <= 1e-8 * abs(db.item())
    assert abs(ft.grad.item() - df.item()) <= 1e-8 * abs(df.item())


def test_sigma8_gradients(engine, bg, layout):
    import torch
    from coffe_amd.autograd import LocalWeights, multipoles, rescale_sigma8
    b, k, p, sep = layout
    lw = LocalWeights(engine, bg, [0.8, 1.6], [0, 2, 4], TABLES, sep, den=1, rsd=1, d1=0)
    A = lw.weights(b=1.5)
    gout = torch.randn((4, 2, 3, len(sep)), dtype=torch.float64, device="cuda",
                       generator=torch.Generator(device="cuda").manual_seed(9))
    s8 = torch.tensor(0.81, dtype=torch.float64, device="cuda", requires_grad=True)
    pk = p.clone().requires_grad_(True)
    L = (multipoles(b, k, rescale_sigma8(k, pk, b.offsets, H, s8), A) * gout).sum()
    L.backward()
    assert abs(s8.grad.item() - 2 * L.item() / 0.81) <= 1e-12 * abs(L.item())
    euler = torch.dot(pk.grad, p).item()
    assert abs(euler) <= 1e-10 * torch.linalg.norm(pk.grad).item() * torch.linalg.norm(p).item()
