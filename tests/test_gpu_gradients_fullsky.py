"""Gradients of the full-sky multipoles at the size of the gradient configuration: one
P(k) with 16384 log-k samples, the 8 default tables, full-sky density + RSD + Doppler-1
at 8 mean redshifts in [0.5, 2], l = 0, 2, 4, 512 separations 5-300 Mpc/h, GL-256 mu
rule.  d/dP is projected on 128 log-k bins (tent functions in ln k times P) and checked
against central finite differences with a relative step of 1e-4 at 32 random
(z, l, r) points, along random directions in the 128-bin space; d/db, d/df (through
the bias and growth-rate polynomial of the weights) and d/dsigma8 at the same points.
"""
import numpy as np
import pytest

from tests.synth import spectrum

pytestmark = pytest.mark.gpu

H = 0.67
TABLES = [(0, 0), (0, 2), (0, 4), (1, 1), (1, 3), (2, 0), (2, 2), (3, 1)]


def test_fullsky_gradients_128_bins(engine):
    import torch
    from coffe_amd import background as B
    from coffe_amd.autograd import LocalWeights, multipoles, rescale_sigma8
    from coffe_amd.engine import Batch
    k_np, p_np = spectrum(16384)
    sep = np.linspace(5.0, 300.0, 512) / H
    zs = list(np.linspace(0.5, 2.0, 8))
    bgd = B.compute(omega_cdm=0.26, omega_baryon=0.05, h=H)
    bg = dict(z=bgd["z"], chi=bgd["chi"], D1=bgd["D1"], f=bgd["f"], cH=bgd["conformal_Hz"], G1=bgd["G1"],
              G2=bgd["G2"])
    print("full-sky gradients: weights", flush=True)
    lw = LocalWeights(engine, bg, zs, [0, 2, 4], TABLES, sep, den=1, rsd=1, d1=1, flatsky=False, f_override=True)
    b = Batch(engine, [len(k_np)], [len(k_np)], [k_np[0]], [k_np[-1]], TABLES, sep, len(zs), 3)
    dev = dict(device="cuda", dtype=torch.float64)
    k = torch.tensor(k_np, **dev)
    p0 = torch.tensor(p_np, **dev)
    # 128 tent functions on ln k
    lk = np.log(k_np)
    nodes = np.linspace(lk[0], lk[-1], 128)
    h_node = nodes[1] - nodes[0]
    phi = np.clip(1.0 - np.abs(lk[None, :] - nodes[:, None]) / h_node, 0.0, None)   # [128][n]
    phi_t = torch.tensor(phi, **dev)

    def pk_of(theta):
        return p0 * (1.0 + theta @ phi_t)

    rng = np.random.default_rng(17)
    pts = [(int(rng.integers(len(zs))), int(rng.integers(3)), int(rng.integers(len(sep)))) for _ in range(32)]
    bv, fv = 1.5, 0.7
    A = lw.weights(b=bv, f=fv)
    theta0 = torch.zeros(128, **dev)
    print("full-sky gradients: 32 adjoint calls", flush=True)
    grads = []
    for z, l, r in pts:
        gout = torch.zeros((1, len(zs), 3, len(sep)), **dev)
        gout[0, z, l, r] = 1.0
        _, _, gpk = b.vjp(k, pk_of(theta0), A, gout)
        grads.append((phi_t @ (gpk * p0)).cpu().numpy())           # dM/dtheta_b
    grads = np.array(grads)                                        # [32][128]
    step = 1e-4
    for trial in range(2):
        u = torch.tensor(rng.standard_normal(128), **dev)
        mp = b.forward(k, pk_of(step * u), A)[1].cpu().numpy()[0]
        mm = b.forward(k, pk_of(-step * u), A)[1].cpu().numpy()[0]
        un = u.cpu().numpy()
        for i, (z, l, r) in enumerate(pts):
            fd = (mp[z, l, r] - mm[z, l, r]) / (2 * step)
            an = grads[i] @ un
            scale = np.max(np.abs(grads[i])) * np.sum(np.abs(un))
            assert abs(an - fd) <= 1e-6 * scale, (trial, i, an, fd)
    # b, f and sigma8 at the same points
    print("full-sky gradients: b, f, sigma8", flush=True)
    sel = torch.zeros((1, len(zs), 3, len(sep)), **dev)
    for z, l, r in pts:
        sel[0, z, l, r] = 1.0
    bt = torch.tensor(bv, **dev, requires_grad=True)
    ft = torch.tensor(fv, **dev, requires_grad=True)
    L = (multipoles(b, k, p0, lw.weights(b=bt, f=ft)) * sel).sum()
    L.backward()
    hb = 1e-4
    with torch.no_grad():
        def loss(x, y):
            return (multipoles(b, k, p0, lw.weights(b=x, f=y)) * sel).sum().item()
        db = (loss(bv * (1 + hb), fv) - loss(bv * (1 - hb), fv)) / (2 * hb * bv)
        df = (loss(bv, fv * (1 + hb)) - loss(bv, fv * (1 - hb))) / (2 * hb * fv)
    assert abs(bt.grad.item() - db) <= 1e-7 * abs(db)
    assert abs(ft.grad.item() - df) <= 1e-7 * abs(df)
    s8 = torch.tensor(0.81, **dev, requires_grad=True)
    L8 = (multipoles(b, k, rescale_sigma8(k, p0, b.offsets, H, s8), A) * sel).sum()
    L8.backward()
    with torch.no_grad():
        def l8(x):
            return (multipoles(b, k, rescale_sigma8(k, p0, b.offsets, H, torch.tensor(x, **dev)), A) * sel).sum().item()
        d8 = (l8(0.81 * (1 + hb)) - l8(0.81 * (1 - hb))) / (2 * hb * 0.81)
    assert abs(s8.grad.item() - d8) <= 1e-7 * abs(d8)
    b.close()


def test_fullsky_vjp_256_spectra(engine):
    """The gradient configuration recomputed over a batch of 256 seeded spectra of 16384
    samples each (full-sky density + RSD + Doppler-1, 8 z, l = 0, 2, 4, 512 separations, 8
    tables) in one adjoint call: per spectrum the homogeneity identity <gpk_s, P_s> =
    <gout_s, M_s> (1e-10); for three spectra the exact dot product against the oracle's
    forward-mode derivative along white-noise directions (1e-9) and central differences
    along random combinations of 128 ln-k tents (relative step 1e-4, 1e-6)."""
    import torch
    import oracle as O
    from coffe_amd import background as B
    from coffe_amd.autograd import LocalWeights
    from coffe_amd.engine import Batch
    from tests.synth import batch_spectra
    specs = [(s[0], s[1]) for s in batch_spectra(256, seed=23, lengths=(16384,))]
    sep = np.linspace(5.0, 300.0, 512) / H
    zs = list(np.linspace(0.5, 2.0, 8))
    bgd = B.compute(omega_cdm=0.26, omega_baryon=0.05, h=H)
    bg = dict(z=bgd["z"], chi=bgd["chi"], D1=bgd["D1"], f=bgd["f"], cH=bgd["conformal_Hz"], G1=bgd["G1"],
              G2=bgd["G2"])
    print("256-spectrum full-sky vjp: weights", flush=True)
    lw = LocalWeights(engine, bg, zs, [0, 2, 4], TABLES, sep, den=1, rsd=1, d1=1, flatsky=False)
    A = lw.weights(b=1.5)
    n = [len(s[0]) for s in specs]
    b = Batch(engine, n, n, [s[0][0] for s in specs], [s[0][-1] for s in specs], TABLES, sep, len(zs), 3)
    dev = dict(device="cuda", dtype=torch.float64)
    k = torch.tensor(np.concatenate([s[0] for s in specs]), **dev)
    p_np = np.concatenate([s[1] for s in specs])
    p = torch.tensor(p_np, **dev)
    rng = np.random.default_rng(29)
    gout = torch.tensor(rng.standard_normal((256, len(zs), 3, len(sep))), **dev)
    print("256-spectrum full-sky vjp: adjoint", flush=True)
    _, out, gpk = b.vjp(k, p, A, gout)
    off = np.cumsum([0] + n)
    gpk_np = gpk.cpu().numpy()
    lhs = np.array([gpk_np[off[s]:off[s + 1]] @ p_np[off[s]:off[s + 1]] for s in range(256)])
    rhs = (gout * out).sum(dim=(1, 2, 3)).cpu().numpy()
    scale = torch.abs(gout * out).sum(dim=(1, 2, 3)).cpu().numpy()
    worst = float(np.max(np.abs(lhs - rhs) / scale))
    print(f"  homogeneity over 256 spectra: worst {worst:.2e}", flush=True)
    assert worst <= 1e-10
    A_np = A.cpu().numpy()
    gout_np = gout.cpu().numpy()
    for s in (0, 117, 255):
        ks, ps = specs[s]
        v = ps * rng.standard_normal(len(ps))
        _, tang = O.tables_at_sep_jvp([(ks, ps)], [len(ks)], TABLES, sep, [v])
        Jv = np.einsum("zltr,tr->zlr", A_np, tang[0])
        l1 = float(np.sum(gout_np[s] * Jv))
        r1 = float(gpk_np[off[s]:off[s + 1]] @ v)
        sc = float(np.sum(np.abs(gout_np[s] * Jv))) + float(np.sum(np.abs(gpk_np[off[s]:off[s + 1]] * v)))
        print(f"  spectrum {s}: <gout, Jv> {l1:.15e} <gpk, v> {r1:.15e} rel {abs(l1 - r1) / sc:.2e}", flush=True)
        assert abs(l1 - r1) <= 1e-9 * sc, (s, l1, r1)
        # 128 tents in ln k, relative step 1e-4
        lk = np.log(ks)
        nodes = np.linspace(lk[0], lk[-1], 128)
        phi = np.clip(1.0 - np.abs(lk[None, :] - nodes[:, None]) / (nodes[1] - nodes[0]), 0.0, None)
        u = rng.standard_normal(128)
        dv = ps * (u @ phi)
        step = 1e-4
        pp, pm = p_np.copy(), p_np.copy()
        pp[off[s]:off[s + 1]] += step * dv
        pm[off[s]:off[s + 1]] -= step * dv
        mp = b.forward(k, torch.tensor(pp, **dev), A)[1][s]
        mm = b.forward(k, torch.tensor(pm, **dev), A)[1][s]
        fd = float((gout[s] * (mp - mm)).sum().item()) / (2 * step)
        an = float(gpk_np[off[s]:off[s + 1]] @ dv)
        sc2 = float(np.sum(np.abs(gpk_np[off[s]:off[s + 1]] * dv)))
        print(f"  spectrum {s}: tents adjoint {an:.15e} fd {fd:.15e}", flush=True)
        assert abs(an - fd) <= 1e-6 * sc2, (s, an, fd)
    b.close()
