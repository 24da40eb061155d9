"""Differentiable multipoles and covariance: torch.autograd wrappers over the batched
engine and the 2D FFTLog covariance (CovarianceFunction: coffe_cov_assemble_vjp and
coffe_cov_integrals_vjp, gradients for P, the biases and f of the z bins).

The multipole of spectrum s at (z, l, r) is linear in the tables and in the weights,
    M[s, z, l, r] = sum_t A[z, l, t, r] I_t[s, r](P_s),
and the tables are a (nonlinear, through Akima) function of the P(k) samples.  The
gradient with respect to P is the engine's adjoint (coffe_batch_vjp); the gradient
with respect to the weights is sum_s gM[s] I[s].

Derivative parameters (the reference has no such API; its galaxy bias and growth
rate enter the local terms of functions.c:129-1143 as):
- b: a constant galaxy bias of both tracers.  den-den ~ b^2, den-rsd and den-d1 ~ b.
- f: a constant growth rate overriding f(z) in every term.  rsd-rsd, d1-d1 and
  rsd-d1 ~ f^2; den-rsd and den-d1 ~ f.
- sigma8: the input spectrum is rescaled to P sigma8^2 / sigma8(P)^2, with
  sigma8(P)^2 = 1/(2 pi^2) int dk k^2 P(k) W(k R)^2, W(x) = 3 (sin x - x cos x) / x^3,
  R = 8 Mpc/h (trapezoid in ln k over the input samples).  The reference applies no
  sigma8 to a user-supplied P(k) (coffe.pyx:2929-2996).
So the weights are the polynomial A = b^2 A_bb + b f A_bf + f^2 A_ff, whose three
groups the engine computes once (coffe_mu_weights_terms); autograd differentiates
the polynomial.
"""
import numpy as np
import torch

from .engine import (TERM_D1_D1, TERM_DEN_D1, TERM_DEN_DEN, TERM_DEN_RSD, TERM_RSD_D1, TERM_RSD_RSD, mu_weights)

GROUP_BB = TERM_DEN_DEN
GROUP_BF = TERM_DEN_RSD | TERM_DEN_D1
GROUP_FF = TERM_RSD_RSD | TERM_D1_D1 | TERM_RSD_D1


class MultipolesFunction(torch.autograd.Function):
    """M = batch(P, A) with gradients for P (adjoint kernels) and A (sum over spectra)."""

    @staticmethod
    def forward(ctx, pk, A, batch, k):
        itab, out = batch.forward(k, pk, A)
        ctx.save_for_backward(pk, A, itab)
        ctx.batch, ctx.k = batch, k
        return out

    @staticmethod
    def backward(ctx, gout):
        pk, A, itab = ctx.saved_tensors
        gout = gout.contiguous()
        gpk = gA = None
        if ctx.needs_input_grad[0]:
            _, _, gpk = ctx.batch.vjp(ctx.k, pk, A, gout)
        if ctx.needs_input_grad[1]:
            gA = torch.einsum("szlr,str->zltr", gout, itab)
        return gpk, gA, None, None


def multipoles(batch, k, pk, A):
    """Differentiable batched multipoles (nspec, nz, nl, nsep) of the concatenated spectra."""
    return MultipolesFunction.apply(pk, A, batch, k)


class CovarianceFunction(torch.autograd.Function):
    """Gaussian covariance cov = assemble(integrals(P); bias, f) with gradients for P
    (coffe_cov_assemble_vjp, then coffe_cov_integrals_vjp), the biases and f (the
    coefficient polynomials of the assembly, summed over the spectra)."""

    @staticmethod
    def forward(ctx, pk, bias, f, plan, k, n_in, zargs, flags):
        bias_np = bias.detach().cpu().numpy()
        f_np = f.detach().cpu().numpy()
        integ = plan.integrals(k, pk, n_in)
        cov = plan.assemble(integ, bias_np, f_np, **zargs, **flags)
        ctx.save_for_backward(pk, integ)
        ctx.plan, ctx.k, ctx.n_in, ctx.zargs, ctx.flags = plan, k, n_in, zargs, flags
        ctx.bias_np, ctx.f_np, ctx.bias_shape = bias_np, f_np, bias.shape
        return cov

    @staticmethod
    def backward(ctx, gout):
        pk, integ = ctx.saved_tensors
        z = ctx.zargs
        ginteg, gbias, gf = ctx.plan.assemble_vjp(integ, gout.contiguous(), ctx.bias_np, ctx.f_np, z["D1"],
                                                  z["nbar"], z["vol"], mixed=ctx.flags["mixed"],
                                                  cosmic=ctx.flags["cosmic"], pops=ctx.flags["pops"],
                                                  nbar2=z["nbar2"])
        gpk = gb = gfr = None
        if ctx.needs_input_grad[0]:
            gpk = ctx.plan.integrals_vjp(ctx.k, pk, ctx.n_in, ginteg)
        if ctx.needs_input_grad[1]:
            g = gbias.sum(axis=0)
            if len(ctx.bias_shape) == 1:      # one population: the four biases are the one b
                g = g.sum(axis=1)
            gb = torch.tensor(g, dtype=pk.dtype)
        if ctx.needs_input_grad[2]:
            gfr = torch.tensor(gf.sum(axis=0), dtype=pk.dtype)
        return gpk, gb, gfr, None, None, None, None, None


def covariance(plan, k, pk, n_in, bias, f, D1, nbar, vol, pix, poisson=True, mixed=True, cosmic=True, pops=None,
               nbar2=None, window=False):
    """Differentiable Gaussian covariance (nspec, nz, nl, R, nl, R) of the concatenated
    spectra (k, pk on the GPU, n_in samples each); bias is a CPU tensor (nz, 4) (galaxy
    biases 1..4, read through the populations pops, default 1, 1, 1, 1) or (nz,) (one value
    in all four), f a CPU tensor (nz,); nbar / nbar2 the densities of the first and second
    tracer.  Gradients flow to pk (and through rescale_sigma8 to sigma8), bias and f."""
    zargs = dict(D1=D1, nbar=nbar, nbar2=nbar if nbar2 is None else nbar2, vol=vol, pix=pix)
    flags = dict(poisson=poisson, mixed=mixed, cosmic=cosmic, pops=pops, window=window)
    return CovarianceFunction.apply(pk, bias, f, plan, k, n_in, zargs, flags)


def sigma8_of(k, pk, offsets, h):
    """sigma8 of each concatenated spectrum (k in 1/Mpc, P in Mpc^3), as a torch tensor."""
    R = 8.0 / h
    out = []
    for s in range(len(offsets) - 1):
        ks, ps = k[offsets[s]:offsets[s + 1]], pk[offsets[s]:offsets[s + 1]]
        x = ks * R
        w = 3 * (torch.sin(x) - x * torch.cos(x)) / x ** 3
        y = ks ** 3 * ps * w * w / (2 * np.pi ** 2)
        out.append(torch.sqrt(torch.trapezoid(y, torch.log(ks))))
    return torch.stack(out)


def rescale_sigma8(k, pk, offsets, h, sigma8):
    """P sigma8^2 / sigma8(P)^2 per spectrum; sigma8 is a scalar or a (nspec,) tensor."""
    s8 = sigma8_of(k, pk, offsets, h)
    scale = (torch.as_tensor(sigma8, dtype=pk.dtype, device=pk.device) / s8) ** 2
    counts = torch.as_tensor(np.diff(offsets), device=pk.device)
    return pk * torch.repeat_interleave(scale.expand(len(counts)), counts)


class GrowthAt(torch.autograd.Function):
    """(D1(z), f(z)) at the mean redshifts as functions of the six background parameters
    theta = (Omega_cdm, Omega_baryon, Omega_gamma, w0, wa, h) (a CPU tensor): values from
    coffe_background_compute, the Jacobian from coffe_background_jvp (one direction per
    parameter); both are read off the tables by Akima at z (the tangent as the Akima of
    the tangent table, which on the smooth 10000-bin tables differs from the derivative
    of the interpolant at the interpolation-error level)."""

    @staticmethod
    def forward(ctx, theta, zmean, bins):
        from . import background as B
        from .engine import akima
        th = theta.detach().cpu().numpy().astype(np.float64)
        kw = dict(zip(B.PARAMS, th.tolist()), bins=bins)
        zq = np.asarray(zmean, dtype=np.float64)
        jac = np.zeros((len(zq), 2, len(B.PARAMS)))
        tab = None
        for q, name in enumerate(B.PARAMS):
            tab, der = B.jvp({name: 1.0}, **kw)
            jac[:, 0, q] = akima(tab["z"], der["D1"], zq)
            jac[:, 1, q] = akima(tab["z"], der["f"], zq)
        out = np.stack([akima(tab["z"], tab["D1"], zq), akima(tab["z"], tab["f"], zq)], axis=1)
        ctx.jac = torch.tensor(jac, dtype=theta.dtype)
        return torch.tensor(out, dtype=theta.dtype)

    @staticmethod
    def backward(ctx, gout):
        return torch.einsum("zc,zcq->q", gout.to(ctx.jac.dtype).cpu(), ctx.jac), None, None


def growth_at(theta, zmean, bins=10000):
    """Differentiable (nz, 2) tensor [D1(z), f(z)] at the mean redshifts; theta is a CPU
    tensor (Omega_cdm, Omega_baryon, Omega_gamma, w0, wa, h)."""
    return GrowthAt.apply(theta, list(zmean), bins)


def flat_weights_of_background(lw, theta, theta0, zmean, b, bins=10000):
    """Flat-sky density + RSD weights with the growth taken from the background at theta:
    A(theta) = (D1(z; theta) / D1(z; theta0))^2 A(b, f(z; theta)), where lw are the
    f-override flat-sky LocalWeights built on the background at theta0 (flat sky, the
    local terms depend on the background only through D1(z)^2 and f(z), functions.c:
    1457-1460).  Differentiable in theta (self-consistent d/dOmega, d/dw, ...) and b."""
    if not lw.f_override:
        raise ValueError("the weights must be built with f_override=True")
    # only flat-sky density + RSD weights with a constant bias depend on the background
    # through D1(z)^2 and f(z) alone; Doppler terms read H, G1 and chi, full-sky terms
    # chi and D1 at the two end points, and bias tables would be multiplied in again
    if not lw.flatsky:
        raise ValueError("flat_weights_of_background needs flat-sky weights (flatsky=True)")
    if lw.d1:
        raise ValueError("flat_weights_of_background does not cover the Doppler terms (d1=False)")
    if not lw.constant_bias:
        raise ValueError("flat_weights_of_background needs constant-bias weights (bias='constant')")
    g = growth_at(theta, zmean, bins)
    g0 = growth_at(theta0.detach(), zmean, bins)
    dev = lw.groups[next(iter(lw.groups))].device
    scale = ((g[:, 0] / g0[:, 0]) ** 2).to(dev)
    f = g[:, 1].to(dev)
    A = None
    for name, coef in (("bb", lambda: b * b * torch.ones_like(f)), ("bf", lambda: b * f), ("ff", lambda: f * f)):
        if name in lw.groups:
            term = (coef() * scale)[:, None, None, None] * lw.groups[name]
            A = term if A is None else A + term
    return A


class LocalWeights:
    """The three weight groups (b^2, b f, f^2) of the local terms for one geometry.

    bg: background tables (z, chi, D1, f, cH, G1, G2).  bias="constant" sets both
    galaxy biases to 1 in the groups and scales them by b in weights(); bias=(bias_z,
    bias1, bias2) keeps those tables and weights() takes no b.  With f_override the
    growth rate is set to 1 in every term and scaled by f in weights().
    """

    def __init__(self, engine, bg, zmean, ls, tables, sep, den=True, rsd=True, d1=False, flatsky=False, nmu=256,
                 bias="constant", f_override=False):
        self.tables = list(tables)
        self.constant_bias = bias is not None
        self.f_override = bool(f_override)
        self.flatsky = bool(flatsky)
        self.d1 = bool(d1)
        bgw = dict(bg)
        if f_override:
            bgw["f"] = np.ones_like(np.asarray(bg["f"]))
        zb = np.linspace(float(np.min(bg["z"])), float(np.max(bg["z"])), 11)
        if isinstance(bias, str):
            if bias != "constant":
                raise ValueError(f"unknown bias mode {bias!r}")
            bz, b1, b2 = zb, np.ones_like(zb), np.ones_like(zb)
        else:
            bz, b1, b2 = bias
            self.constant_bias = False
        mask = 0
        mask |= TERM_DEN_DEN if den else 0
        mask |= TERM_RSD_RSD if rsd else 0
        mask |= TERM_D1_D1 if d1 else 0
        mask |= TERM_DEN_RSD if den and rsd else 0
        mask |= TERM_DEN_D1 if den and d1 else 0
        mask |= TERM_RSD_D1 if rsd and d1 else 0
        self.groups = {}
        for name, g in (("bb", GROUP_BB), ("bf", GROUP_BF), ("ff", GROUP_FF)):
            if mask & g:
                self.groups[name] = mu_weights(engine, bgw, bz, b1, b2, zmean, ls, self.tables, sep,
                                               flatsky=flatsky, nmu=nmu, terms=mask & g)

    def weights(self, b=None, f=None):
        """A(b, f) = b^2 A_bb + b f A_bf + f^2 A_ff as a differentiable torch expression."""
        if self.constant_bias and b is None:
            raise ValueError("constant-bias weights need b")
        if not self.constant_bias and b is not None:
            raise ValueError("b applies to constant-bias weights only")
        if self.f_override and f is None:
            raise ValueError("f-override weights need f")
        if not self.f_override and f is not None:
            raise ValueError("f applies to f-override weights only")
        b = 1.0 if b is None else b
        f = 1.0 if f is None else f
        A = None
        for name, coef in (("bb", lambda: b * b), ("bf", lambda: b * f), ("ff", lambda: f * f)):
            if name in self.groups:
                term = coef() * self.groups[name]
                A = term if A is None else A + term
        return A
