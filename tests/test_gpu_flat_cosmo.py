"""Flat-sky density + RSD multipoles with the growth taken from the background
(coffe_amd.autograd.flat_weights_of_background, growth_at; coffe_background_jvp):
- at theta0 the weights equal those computed by the engine from the background's own
  f(z) (LocalWeights without the f override), 1e-12;
- d<G, M>/d Omega_cdm, d/d w0 and d/d Omega_baryon through autograd against central
  differences of the whole pipeline (background recomputed at theta +- h, weights by the
  engine from it, multipoles), relative step 1e-4, to 1e-6."""
import numpy as np
import pytest

from tests.synth import batch_spectra

pytestmark = pytest.mark.gpu

TABLES = [(0, 0), (0, 2), (0, 4)]
ZS = [0.8, 1.6]


def _bg(theta):
    from coffe_amd import background as B
    b = B.compute(**dict(zip(B.PARAMS, [float(x) for x in theta])))
    return dict(z=b["z"], chi=b["chi"], D1=b["D1"], f=b["f"], cH=b["conformal_Hz"], G1=b["G1"], G2=b["G2"])


def test_flat_growth_from_background(engine):
    import torch
    from coffe_amd.autograd import LocalWeights, flat_weights_of_background, multipoles
    from coffe_amd.engine import Batch
    theta0 = torch.tensor([0.26, 0.049, 9e-5, -0.95, 0.05, 0.68], dtype=torch.float64)
    h = float(theta0[5])
    specs = batch_spectra(3, seed=21, lengths=(1024, 2048))
    sep = np.linspace(10, 250, 16) / h
    bt = Batch(engine, [len(s[0]) for s in specs], [len(s[0]) for s in specs], [s[0][0] for s in specs],
               [s[0][-1] for s in specs], TABLES, sep, len(ZS), 3)
    k = torch.tensor(np.concatenate([s[0] for s in specs]), device="cuda")
    p = torch.tensor(np.concatenate([s[1] for s in specs]), device="cuda")
    bg0 = _bg(theta0)
    print("flat growth: weights at theta0", flush=True)
    lw = LocalWeights(engine, bg0, ZS, [0, 2, 4], TABLES, sep, den=1, rsd=1, flatsky=True, f_override=True)
    bval = 1.4
    A = flat_weights_of_background(lw, theta0.clone(), theta0, ZS, bval)
    direct = LocalWeights(engine, bg0, ZS, [0, 2, 4], TABLES, sep, den=1, rsd=1, flatsky=True).weights(b=bval)
    assert torch.max(torch.abs(A - direct)).item() <= 1e-12 * torch.max(torch.abs(direct)).item()
    G = torch.tensor(np.random.default_rng(5).standard_normal((3, len(ZS), 3, len(sep))), dtype=torch.float64,
                     device="cuda")
    th = theta0.clone().requires_grad_(True)
    L = (multipoles(bt, k, p, flat_weights_of_background(lw, th, theta0, ZS, bval)) * G).sum()
    L.backward()

    def loss(theta):
        w = LocalWeights(engine, _bg(theta), ZS, [0, 2, 4], TABLES, sep, den=1, rsd=1, flatsky=True).weights(b=bval)
        return (multipoles(bt, k, p, w) * G).sum().item()
    for q in (0, 1, 3):
        x0 = float(theta0[q])
        step = 1e-4 * abs(x0)
        tp, tm = theta0.clone(), theta0.clone()
        tp[q] += step
        tm[q] -= step
        fd = (loss(tp) - loss(tm)) / (2 * step)
        print(f"  d/dtheta[{q}]: autograd {th.grad[q].item():.12e} fd {fd:.12e}", flush=True)
        assert abs(th.grad[q].item() - fd) <= 1e-6 * abs(fd), (q, th.grad[q].item(), fd)
    bt.close()
