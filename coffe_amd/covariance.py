"""Python handle on the 2D FFTLog covariance (include/coffe_twobessel.h)."""
import ctypes

import numpy as np

from ._lib import check, load_library
from .engine import _need, _ptr


class Config(ctypes.Structure):
    """The reference's config of two_sph_bessel (twobessel.h:1-6)."""
    _fields_ = [("l1", ctypes.c_double), ("l2", ctypes.c_double), ("nu1", ctypes.c_double),
                ("nu2", ctypes.c_double), ("c_window_width", ctypes.c_double), ("Nk_sample", ctypes.c_long)]


def two_sph_bessel(k1, k2, fk, l1, l2, nu1=1.01, nu2=1.01, width=0.25):
    """Drop-in two_sph_bessel (twobessel.c:16-113) on host arrays: returns r1, r2, result."""
    L = load_library()
    k1 = np.ascontiguousarray(k1, dtype=np.float64)
    k2 = np.ascontiguousarray(k2, dtype=np.float64)
    fk = np.ascontiguousarray(fk, dtype=np.float64)
    N1, N2 = len(k1), len(k2)
    if fk.shape != (N1, N2):
        raise ValueError("fk must be N1 x N2")
    r1, r2, res = np.empty(N1), np.empty(N2), np.empty((N1, N2))
    dp = ctypes.POINTER(ctypes.c_double)
    rows_in = (dp * N1)(*[fk[i].ctypes.data_as(dp) for i in range(N1)])
    rows_out = (dp * N1)(*[res[i].ctypes.data_as(dp) for i in range(N1)])
    cfg = Config(l1, l2, nu1, nu2, width, N1)
    check(L.two_sph_bessel_status(k1.ctypes.data_as(dp), k2.ctypes.data_as(dp), rows_in, N1, N2,
                                  ctypes.byref(cfg), r1.ctypes.data_as(dp), r2.ctypes.data_as(dp), rows_out),
          "two_sph_bessel")
    return r1, r2, res


class CovPlan:
    """Spectrum-independent tables of the FFTLog covariance for (N, kmin, kmax, ls, sep)."""

    def __init__(self, engine, N, kmin, kmax, ls, sep):
        self.L = engine.L
        self.ls = list(ls)
        self.sep = np.ascontiguousarray(sep, dtype=np.float64)
        self.R = len(self.sep)
        lsa = (ctypes.c_int * len(self.ls))(*self.ls)
        h = ctypes.c_void_p()
        check(self.L.coffe_cov_plan_create(engine.h, int(N), float(kmin), float(kmax), len(self.ls), lsa, self.R,
                                           self.sep.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                           ctypes.byref(h)), "coffe_cov_plan_create")
        self.h = h
        self.npair = self.L.coffe_cov_plan_pairs(h)
        self.nrows = self.L.coffe_cov_plan_rows(h)
        self.pairs = [(i, j) for i in range(len(self.ls)) for j in range(i, len(self.ls))]

    def close(self):
        if self.h:
            self.L.coffe_cov_plan_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def integrals(self, k, pk, n_in, out=None, stream=None):
        """(nspec, 2, npair, R, R) tensor; [s, p-1, pair, n, m] is the value at (sep_m, sep_n)."""
        import torch
        n_in = np.ascontiguousarray(n_in, dtype=np.int64)
        total = int(n_in.sum())
        _need(k, (total,), "k")
        _need(pk, (total,), "pk")
        nspec = len(n_in)
        if out is None:
            out = torch.empty((nspec, 2, self.npair, self.R, self.R), dtype=torch.float64, device=k.device)
        _need(out, (nspec, 2, self.npair, self.R, self.R), "out")
        st = stream if stream is not None else ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        check(self.L.coffe_cov_integrals(self.h, nspec, n_in.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                         _ptr(k), _ptr(pk), _ptr(out), st), "coffe_cov_integrals")
        return out

    def integrals_vjp(self, k, pk, n_in, gint, gpk=None, stream=None):
        """Vector-Jacobian product of `integrals`: the gradient of sum(gint * integrals(k, pk))
        with respect to pk (the Akima interpolation of the samples included).  gint has the
        shape of `integrals`' output; returns gpk, shaped like pk."""
        import torch
        n_in = np.ascontiguousarray(n_in, dtype=np.int64)
        total = int(n_in.sum())
        _need(k, (total,), "k")
        _need(pk, (total,), "pk")
        nspec = len(n_in)
        _need(gint, (nspec, 2, self.npair, self.R, self.R), "gint")
        if gpk is None:
            gpk = torch.empty_like(pk)
        _need(gpk, (total,), "gpk")
        st = stream if stream is not None else ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        check(self.L.coffe_cov_integrals_vjp(self.h, nspec, n_in.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                             _ptr(k), _ptr(pk), _ptr(gint), _ptr(gpk), st),
              "coffe_cov_integrals_vjp")
        return gpk

    @staticmethod
    def _pops(pops):
        p = (1, 1, 1, 1) if pops is None else tuple(int(v) for v in pops)
        if len(p) != 4:
            raise ValueError("pops must have length 4")
        return (ctypes.c_int * 4)(*p)

    def _bin_args(self, bias, per_bin):
        f64 = lambda a: np.ascontiguousarray(np.atleast_1d(a), dtype=np.float64)  # noqa: E731
        bias = np.asarray(bias, dtype=np.float64)
        if bias.ndim == 1:
            bias = np.repeat(bias[:, None], 4, axis=1)
        if bias.ndim != 2 or bias.shape[1] != 4:
            raise ValueError("bias must be (nz,) or (nz, 4)")
        bias = f64(bias)
        nz = bias.shape[0]
        arrs = [f64(a) for a in per_bin]
        if any(len(a) != nz for a in arrs):
            raise ValueError("per-bin arrays must have one entry per z bin")
        return bias, nz, arrs

    def assemble(self, integ, bias, f, D1, nbar, vol, pix, poisson=True, mixed=True, cosmic=True, pops=None,
                 nbar2=None, window=False, stream=None):
        """Gaussian covariance (nspec, nz, nl, R, nl, R) (covariance.c:790-1060).  bias is (nz, 4),
        the galaxy biases 1..4 per bin, or (nz,) for one value in all four; pops are the
        populations covariance_pop1..4 (default 1, 1, 1, 1); nbar / nbar2 the densities of the
        first and second tracer (nbar2 defaults to nbar); window the bin-averaged Poisson term."""
        import torch
        nspec = integ.shape[0]
        _need(integ, (nspec, 2, self.npair, self.R, self.R), "integ")
        bias, nz, arrs = self._bin_args(bias, (f, D1, nbar, nbar if nbar2 is None else nbar2, vol, pix))
        nl = len(self.ls)
        out = torch.empty((nspec, nz, nl, self.R, nl, self.R), dtype=torch.float64, device=integ.device)
        dp = ctypes.POINTER(ctypes.c_double)
        st = stream if stream is not None else ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        check(self.L.coffe_cov_assemble(self.h, nspec, nz, self._pops(pops), bias.ctypes.data_as(dp),
                                        *[a.ctypes.data_as(dp) for a in arrs], int(poisson), int(mixed), int(cosmic),
                                        int(window), _ptr(integ), _ptr(out), st), "coffe_cov_assemble")
        return out

    def assemble_vjp(self, integ, gout, bias, f, D1, nbar, vol, mixed=True, cosmic=True, pops=None, nbar2=None,
                     stream=None):
        """Vector-Jacobian product of `assemble` for the cotangent gout (its output's shape):
        (ginteg shaped like integ, gbias (nspec, nz, 4), gf (nspec, nz)), the gradients of
        sum(gout * cov) per spectrum with respect to the integrals, the galaxy biases 1..4 and f
        of every z bin (the Poisson term depends on none of them)."""
        import torch
        nspec = integ.shape[0]
        _need(integ, (nspec, 2, self.npair, self.R, self.R), "integ")
        bias, nz, arrs = self._bin_args(bias, (f, D1, nbar, nbar if nbar2 is None else nbar2, vol))
        nl = len(self.ls)
        _need(gout, (nspec, nz, nl, self.R, nl, self.R), "gout")
        ginteg = torch.empty_like(integ)
        gbias = np.zeros((nspec, nz, 4))
        gf = np.zeros((nspec, nz))
        dp = ctypes.POINTER(ctypes.c_double)
        st = stream if stream is not None else ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        check(self.L.coffe_cov_assemble_vjp(self.h, nspec, nz, self._pops(pops), bias.ctypes.data_as(dp),
                                            *[a.ctypes.data_as(dp) for a in arrs], int(mixed), int(cosmic),
                                            _ptr(integ), _ptr(gout), _ptr(ginteg), gbias.ctypes.data_as(dp),
                                            gf.ctypes.data_as(dp), st), "coffe_cov_assemble_vjp")
        return ginteg, gbias, gf
