"""Gradients of the Gaussian covariance (coffe_cov_assemble_vjp, cov_kernels.hip
k_cov_assemble_T / k_cov_coef_T, with coffe_cov_integrals_vjp; coffe_amd.autograd.covariance)
with respect to the integrals, the galaxy biases and f of every z bin, P(k) and sigma8:
- the integrals: the covariance without its Poisson term is linear in them, so
  <ginteg, X> = <G, assemble(X)> for any X (1e-12), for X random and X = the integrals;
- the biases and f: central differences of <G, cov> (relative step 1e-4; cov is a
  polynomial of degree 4 in them) to 1e-7;
- P: the mixed term is homogeneous of degree 1 in P and the cosmic one of degree 2, so
  sum gpk P = <G, cov_mixed> + 2 <G, cov_cosmic> (1e-11);
- sigma8 (P rescaled to sigma8^2 / sigma8(P)^2): d<G, cov>/d sigma8 =
  (2 <G, cov_mixed> + 4 <G, cov_cosmic>) / sigma8 (1e-10)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
H = 0.67


def _setup(engine):
    import torch
    from coffe_amd.covariance import CovPlan
    from tests.synth import batch_spectra
    specs = batch_spectra(2, seed=6, lengths=(1024, 2048))
    kmin = max(s[0][0] for s in specs)
    kmax = min(s[0][-1] for s in specs)
    sep = np.array([40.0, 75.0, 120.0, 180.0, 250.0])
    plan = CovPlan(engine, 1024, kmin, kmax, [0, 2, 4], sep)
    dev = dict(device="cuda", dtype=torch.float64)
    k = torch.tensor(np.concatenate([s[0] for s in specs]), **dev)
    p = torch.tensor(np.concatenate([s[1] for s in specs]), **dev)
    n_in = [len(s[0]) for s in specs]
    zargs = dict(D1=[0.8, 0.6], nbar=[1e-3 * H ** 3, 4e-4 * H ** 3], vol=[2e9 / H ** 3, 5e9 / H ** 3],
                 pix=[5.0 / H, 5.0 / H])
    bias = np.array([[1.4, 1.6, 1.5, 1.3], [2.0, 2.2, 1.9, 2.1]])
    f = np.array([0.75, 0.86])
    return plan, k, p, n_in, zargs, bias, f, dev


def _dot(a, b):
    return float((a * b).sum().item())


@pytest.mark.parametrize("pops", [None, (1, 2, 1, 2), (2, 1, 1, 2), (1, 3, 4, 2)])
def test_assemble_vjp(engine, pops):
    """Default populations (only galaxy bias 1 enters: the others' gradients are zero) and
    multi-tracer populations with density2 != density1."""
    import torch
    plan, k, p, n_in, z, bias, f, dev = _setup(engine)
    print(f"cov assemble vjp: rows={plan.nrows} pops={pops}", flush=True)
    z = dict(z, pops=pops, nbar2=[2.5e-3 * H ** 3, 7e-4 * H ** 3])
    integ = plan.integrals(k, p, n_in)
    cov = plan.assemble(integ, bias, f, **z)
    G = torch.tensor(np.random.default_rng(3).standard_normal(tuple(cov.shape)), **dev)
    zv = dict(D1=z["D1"], nbar=z["nbar"], vol=z["vol"], pops=pops, nbar2=z["nbar2"])
    ginteg, gbias, gf = plan.assemble_vjp(integ, G, bias, f, **zv)
    used = {0} if pops is None else {q - 1 for q in pops}
    for q in range(4):
        if q not in used:
            assert np.all(gbias[:, :, q] == 0.0), (q, gbias[:, :, q])
    # linearity in the integrals (no Poisson term)
    X = torch.tensor(np.random.default_rng(4).standard_normal(tuple(integ.shape)), **dev)
    for Y in (X, integ):
        lhs = _dot(ginteg, Y)
        cy = plan.assemble(Y, bias, f, **z, poisson=False)
        rhs = _dot(G, cy)
        ref = float(torch.abs(G * cy).sum().item())
        print(f"  linearity: {lhs:.15e} {rhs:.15e}", flush=True)
        assert abs(lhs - rhs) <= 1e-12 * ref
    # biases and f by central differences, per spectrum and z bin
    step = 1e-4

    def L(b, ff):
        c = plan.assemble(integ, b, ff, **z)
        return (G * c).sum(dim=(1, 2, 3, 4, 5)).cpu().numpy()          # per spectrum
    for zb in range(2):
        for q in range(5):
            bp, bm, fp, fm = bias.copy(), bias.copy(), f.copy(), f.copy()
            if q < 4:
                h = step * bias[zb, q]
                bp[zb, q] += h
                bm[zb, q] -= h
                an = gbias[:, zb, q]
            else:
                h = step * f[zb]
                fp[zb] += h
                fm[zb] -= h
                an = gf[:, zb]
            fd = (L(bp, fp) - L(bm, fm)) / (2 * h)
            print(f"  z{zb} param {q}: adjoint {an} fd {fd}", flush=True)
            assert np.all(np.abs(an - fd) <= 1e-7 * np.abs(fd) + 1e-12 * np.max(np.abs(fd))), (zb, q, an, fd)
    plan.close()


def test_covariance_autograd(engine):
    import torch
    from coffe_amd.autograd import covariance, rescale_sigma8
    plan, k, p, n_in, z, bias, f, dev = _setup(engine)
    G = torch.tensor(np.random.default_rng(8).standard_normal((2, 2, 3, plan.R, 3, plan.R)), **dev)
    print("cov autograd: P, b (one population), f", flush=True)
    pk = p.clone().requires_grad_(True)
    b1 = torch.tensor([1.5, 2.0], dtype=torch.float64, requires_grad=True)
    ft = torch.tensor(f, dtype=torch.float64, requires_grad=True)
    loss = (covariance(plan, k, pk, n_in, b1, ft, **z) * G).sum()
    loss.backward()
    with torch.no_grad():
        cm = covariance(plan, k, p, n_in, b1, ft, **z, poisson=False, cosmic=False)
        cc = covariance(plan, k, p, n_in, b1, ft, **z, poisson=False, mixed=False)
        Lm, Lc = _dot(G, cm), _dot(G, cc)
    lhs = float((pk.grad * p).sum().item())
    ref = float(torch.abs(G * cm).sum().item() + 2 * torch.abs(G * cc).sum().item())
    print(f"  Euler: {lhs:.15e} {Lm + 2 * Lc:.15e}", flush=True)
    assert abs(lhs - (Lm + 2 * Lc)) <= 1e-11 * ref
    # one-population b: the sum of the four bias gradients; against central differences
    step = 1e-4
    for zb in range(2):
        def Lb(x):
            bb = b1.detach().clone()
            bb[zb] = x
            with torch.no_grad():
                return _dot(G, covariance(plan, k, p, n_in, bb, ft.detach(), **z))
        x0 = float(b1[zb].detach())
        fd = (Lb(x0 * (1 + step)) - Lb(x0 * (1 - step))) / (2 * step * x0)
        assert abs(b1.grad[zb].item() - fd) <= 1e-7 * abs(fd), (zb, b1.grad[zb].item(), fd)
    # sigma8: closed form from the degrees of the two terms
    print("cov autograd: sigma8", flush=True)
    offsets = np.cumsum([0] + n_in)
    s8 = torch.tensor(0.81, dtype=torch.float64, device="cuda", requires_grad=True)
    ps8 = rescale_sigma8(k, p, offsets, H, s8)
    L8 = (covariance(plan, k, ps8, n_in, b1.detach(), ft.detach(), **z, poisson=False) * G).sum()
    L8.backward()
    with torch.no_grad():
        ps = ps8.detach()
        m8 = _dot(G, covariance(plan, k, ps, n_in, b1.detach(), ft.detach(), **z, poisson=False, cosmic=False))
        c8 = _dot(G, covariance(plan, k, ps, n_in, b1.detach(), ft.detach(), **z, poisson=False, mixed=False))
    want = (2 * m8 + 4 * c8) / 0.81
    print(f"  sigma8: {s8.grad.item():.15e} {want:.15e}", flush=True)
    assert abs(s8.grad.item() - want) <= 1e-10 * (2 * abs(m8) + 4 * abs(c8)) / 0.81
    plan.close()
