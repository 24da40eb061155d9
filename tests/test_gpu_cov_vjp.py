"""Vector-Jacobian product of the covariance integrals (coffe_cov_integrals_vjp,
cov_kernels.hip: the row functionals, row transforms and diagonal transposed, then
the Akima spline's adjoint) against central finite differences of the forward
integrals: for a random cotangent G and smooth random directions v of P,
<gpk, v> = (<G, I(P + h v)> - <G, I(P - h v)>) / 2h, h = 1e-5, to 1e-7 of
sum |gpk v|, on the fixture's spectrum.

Akima's slope weights are |m_{i+1} - m_i| of neighbouring secants, so the map from
the samples to the integrals has kinks where two secants agree; a finite step that
crosses one gives a difference quotient that matches neither one-sided derivative
(a synthetic spectrum with BAO wiggles has such points within 1e-5 relative of its
samples).  The batch test therefore uses checks with no kink in the way:
- homogeneity: Akima is homogeneous of degree 1 in the samples, the P integrals of
  degree 1 and the P^2 integrals of degree 2, so per spectrum
  sum_i gpk_i P_i = <G_1, I_1> + 2 <G_2, I_2> exactly (Euler);
- affine directions v = a + b k: they add b to every secant and leave the weights
  alone, so I_1 is linear and I_2 quadratic along v and the central difference is
  exact up to rounding."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
H = 0.67


def _directions(ks, n, rng):
    """n directions in the space of samples: P times a random combination of 24 tents
    in ln k (smooth, so the perturbed spectra stay resolved by the FFT grid)."""
    lk = np.log(ks)
    nodes = np.linspace(lk[0], lk[-1], 24)
    phi = np.clip(1.0 - np.abs(lk[None, :] - nodes[:, None]) / (nodes[1] - nodes[0]), 0.0, None)
    return [rng.standard_normal(24) @ phi for _ in range(n)]


def _check(plan, ks_list, ps_list, rng, step=1e-5, tol=1e-7):
    import torch
    dev = dict(device="cuda", dtype=torch.float64)
    n_in = [len(x) for x in ks_list]
    k = torch.tensor(np.concatenate(ks_list), **dev)
    p0 = np.concatenate(ps_list)
    p = torch.tensor(p0, **dev)
    shape = (len(n_in), 2, plan.npair, plan.R, plan.R)
    G = torch.tensor(rng.standard_normal(shape), **dev)
    gpk = plan.integrals_vjp(k, p, n_in, G).cpu().numpy()
    assert np.all(np.isfinite(gpk))
    for trial in range(3):
        rel = np.concatenate([d for d in (_directions(ks, 1, rng)[0] for ks in ks_list)])
        v = p0 * rel
        def cd(h):
            ip = plan.integrals(k, torch.tensor(p0 + h * v, **dev), n_in)
            im = plan.integrals(k, torch.tensor(p0 - h * v, **dev), n_in)
            return ((G * ip).sum() - (G * im).sum()).item() / (2 * h)
        fd = cd(step)
        an = float(gpk @ v)
        scale = float(np.sum(np.abs(gpk * v)))
        print(f"  trial {trial}: adjoint {an:.15e} fd {fd:.15e} scale {scale:.3e}", flush=True)
        assert abs(an - fd) <= tol * scale, (trial, an, fd, scale)
    return gpk


def test_cov_vjp_single_spectrum(engine):
    """The fixture's spectrum, N = 1024, l = 0, 2, 4 (6 pairs), 6 separations."""
    import os
    from coffe_amd.covariance import CovPlan
    gold = os.path.join(os.path.dirname(__file__), "golden")
    k, pk = np.loadtxt(os.path.join(gold, "PkL_CLASS.dat")).T
    k, pk = k * H, pk / H ** 3
    sep = np.arange(50, 350, 50) / H
    plan = CovPlan(engine, 1024, k[0], k[-1], [0, 2, 4], sep)
    print(f"cov vjp: N=1024 rows={plan.nrows}", flush=True)
    _check(plan, [k], [pk], np.random.default_rng(5))
    plan.close()


def test_cov_vjp_batch(engine):
    """Three spectra of different lengths in one call, N = 1000 (radix-10 passes),
    l = 0, 2: the homogeneity identity per spectrum, central differences along affine
    directions per spectrum, and the gradient of each spectrum depends on its own
    cotangent only."""
    import torch
    from coffe_amd.covariance import CovPlan
    from tests.synth import batch_spectra
    specs = batch_spectra(3, seed=4, lengths=(1024, 2048))
    kmin = max(s[0][0] for s in specs)
    kmax = min(s[0][-1] for s in specs)
    sep = np.array([30.0, 47.0, 90.0, 150.0, 260.0])
    plan = CovPlan(engine, 1000, kmin, kmax, [0, 2], sep)
    ks = [s[0] for s in specs]
    ps = [s[1] for s in specs]
    print(f"cov vjp: batch of 3, N=1000 rows={plan.nrows}", flush=True)
    dev = dict(device="cuda", dtype=torch.float64)
    n_in = [len(x) for x in ks]
    o = np.cumsum([0] + n_in)
    k = torch.tensor(np.concatenate(ks), **dev)
    p0 = np.concatenate(ps)
    p = torch.tensor(p0, **dev)
    G = torch.tensor(np.random.default_rng(9).standard_normal((3, 2, plan.npair, plan.R, plan.R)), **dev)
    I = plan.integrals(k, p, n_in)
    g_full = plan.integrals_vjp(k, p, n_in, G).cpu().numpy()
    assert np.all(np.isfinite(g_full))
    for s in range(3):
        lhs = float(g_full[o[s]:o[s + 1]] @ p0[o[s]:o[s + 1]])
        rhs = (G[s, 0] * I[s, 0]).sum().item() + 2 * (G[s, 1] * I[s, 1]).sum().item()
        ref = (torch.abs(G[s, 0] * I[s, 0]).sum() + 2 * torch.abs(G[s, 1] * I[s, 1]).sum()).item()
        print(f"  homogeneity s{s}: {lhs:.15e} {rhs:.15e}", flush=True)
        assert abs(lhs - rhs) <= 1e-12 * ref, (s, lhs, rhs)
    rng = np.random.default_rng(3)
    h = 1e-3
    for s in range(3):
        kk = ks[s]
        a, b = rng.standard_normal(2)
        v = np.zeros_like(p0)
        # a few per cent of P near the peak of the spectrum
        v[o[s]:o[s + 1]] = np.max(ps[s]) * 0.05 * (a + b * kk / kk[-1])
        ip = plan.integrals(k, torch.tensor(p0 + h * v, **dev), n_in)
        im = plan.integrals(k, torch.tensor(p0 - h * v, **dev), n_in)
        fd = ((G * ip).sum() - (G * im).sum()).item() / (2 * h)
        an = float(g_full @ v)
        scale = float(np.sum(np.abs(g_full * v)))
        print(f"  affine s{s}: adjoint {an:.15e} fd {fd:.15e} scale {scale:.3e}", flush=True)
        assert abs(an - fd) <= 1e-9 * scale, (s, an, fd, scale)
    # zero cotangent for spectrum 1: its gradient vanishes, the others are unchanged
    G[1] = 0
    g_part = plan.integrals_vjp(k, p, n_in, G).cpu().numpy()
    assert np.all(g_part[o[1]:o[2]] == 0)
    for s in (0, 2):
        np.testing.assert_allclose(g_part[o[s]:o[s + 1]], g_full[o[s]:o[s + 1]], rtol=0,
                                   atol=1e-12 * np.max(np.abs(g_full[o[s]:o[s + 1]])))
    plan.close()


def test_cov_vjp_dense_shuffled_separations(engine):
    """48 separations 10-200 Mpc/h (several per grid cell, so the row windows of the
    functionals are not monotone in the separation) given in a shuffled order, N = 2048:
    the homogeneity identity, and the gradient equals the one of the sorted plan with
    the cotangent permuted alike."""
    import torch
    from coffe_amd.covariance import CovPlan
    from tests.synth import spectrum
    k_np, p_np = spectrum(4096)
    sep = np.linspace(10.0, 200.0, 48) / H
    perm = np.random.default_rng(1).permutation(len(sep))
    dev = dict(device="cuda", dtype=torch.float64)
    k = torch.tensor(k_np, **dev)
    p = torch.tensor(p_np, **dev)
    ps = CovPlan(engine, 2048, k_np[0], k_np[-1], [0, 2], sep)
    pu = CovPlan(engine, 2048, k_np[0], k_np[-1], [0, 2], sep[perm])
    print(f"cov vjp: 48 shuffled seps, N=2048 rows={ps.nrows}", flush=True)
    G = torch.tensor(np.random.default_rng(2).standard_normal((1, 2, ps.npair, ps.R, ps.R)), **dev)
    Gu = G[:, :, :, perm][:, :, :, :, perm].contiguous()
    Iu = pu.integrals(k, p, [len(k_np)])
    gs = ps.integrals_vjp(k, p, [len(k_np)], G).cpu().numpy()
    gu = pu.integrals_vjp(k, p, [len(k_np)], Gu).cpu().numpy()
    np.testing.assert_allclose(gu, gs, rtol=0, atol=1e-11 * np.max(np.abs(gs)))
    lhs = float(gu @ p_np)
    rhs = (Gu[0, 0] * Iu[0, 0]).sum().item() + 2 * (Gu[0, 1] * Iu[0, 1]).sum().item()
    ref = (torch.abs(Gu[0, 0] * Iu[0, 0]).sum() + 2 * torch.abs(Gu[0, 1] * Iu[0, 1]).sum()).item()
    assert abs(lhs - rhs) <= 1e-12 * ref, (lhs, rhs)
    ps.close()
    pu.close()


def _euler_and_affine(plan, k_np, p_np, seed, h=1e-3, affine_tol=1e-9):
    """Per spectrum: the homogeneity identity sum gpk P = <G_1, I_1> + 2 <G_2, I_2> (1e-12)
    and central differences along two affine directions a + b k (exact up to rounding)."""
    import torch
    dev = dict(device="cuda", dtype=torch.float64)
    k = torch.tensor(k_np, **dev)
    p = torch.tensor(p_np, **dev)
    n_in = [len(k_np)]
    G = torch.tensor(np.random.default_rng(seed).standard_normal((1, 2, plan.npair, plan.R, plan.R)), **dev)
    I = plan.integrals(k, p, n_in)
    g = plan.integrals_vjp(k, p, n_in, G).cpu().numpy()
    assert np.all(np.isfinite(g))
    lhs = float(g @ p_np)
    rhs = (G[0, 0] * I[0, 0]).sum().item() + 2 * (G[0, 1] * I[0, 1]).sum().item()
    ref = (torch.abs(G[0, 0] * I[0, 0]).sum() + 2 * torch.abs(G[0, 1] * I[0, 1]).sum()).item()
    print(f"  homogeneity: {lhs:.15e} {rhs:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * ref, (lhs, rhs)
    rng = np.random.default_rng(seed + 1)
    for trial in range(2):
        a, b = rng.standard_normal(2)
        v = np.max(p_np) * 0.05 * (a + b * k_np / k_np[-1])
        ip = plan.integrals(k, torch.tensor(p_np + h * v, **dev), n_in)
        im = plan.integrals(k, torch.tensor(p_np - h * v, **dev), n_in)
        fd = ((G * ip).sum() - (G * im).sum()).item() / (2 * h)
        an = float(g @ v)
        scale = float(np.sum(np.abs(g * v)))
        print(f"  affine {trial}: adjoint {an:.15e} fd {fd:.15e} scale {scale:.3e}", flush=True)
        assert abs(an - fd) <= affine_tol * scale, (trial, an, fd, scale)


def test_cov_vjp_n8000(engine):
    """The adjoint at the covariance's default length N = 8000 (radix-10 transforms, about
    1200 rows), l = 0, 2, 4, 16 separations 10-200 Mpc/h, the fixture spectrum."""
    import os
    from coffe_amd.covariance import CovPlan
    gold = os.path.join(os.path.dirname(__file__), "golden")
    k, pk = np.loadtxt(os.path.join(gold, "PkL_CLASS.dat")).T
    k, pk = k * H, pk / H ** 3
    sep = np.linspace(10.0, 200.0, 16) / H
    plan = CovPlan(engine, 8000, k[0], k[-1], [0, 2, 4], sep)
    print(f"cov vjp: N=8000 rows={plan.nrows}", flush=True)
    _euler_and_affine(plan, k, pk, 21)
    plan.close()


def test_cov_vjp_more_than_256_separations(engine):
    """300 separations (the transposed column functionals run over them in chunks of 256,
    so the second chunk is exercised), N = 512, l = 0, 2."""
    from coffe_amd.covariance import CovPlan
    from tests.synth import spectrum
    k_np, p_np = spectrum(2048)
    sep = np.linspace(10.0, 250.0, 300) / H
    plan = CovPlan(engine, 512, k_np[0], k_np[-1], [0, 2], sep)
    print(f"cov vjp: 300 seps, N=512 rows={plan.nrows}", flush=True)
    _euler_and_affine(plan, k_np, p_np, 31)
    plan.close()
