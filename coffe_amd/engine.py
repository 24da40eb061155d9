"""Python handle on the batched engine (device memory through torch tensors)."""
import ctypes

import numpy as np

from ._lib import COFFE_ENOCONV, check, load_library

TABLE_ORDER = [(0, 0), (0, 2), (0, 4), (1, 1), (1, 3), (2, 0), (2, 2), (3, 1)]  # (n, l), integrals.c:599-608


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _need(t, shape, name):
    """A tensor handed to a kernel must be contiguous float64 on a GPU with exactly this shape."""
    import torch
    if t is None:
        return
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch tensor")
    if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"{name}: needs a contiguous float64 CUDA tensor (got {t.dtype}, {t.device})")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)} != {tuple(shape)}")


class Engine:
    """One engine per device (stream + cache of spectrum-independent grid tables)."""

    def __init__(self, device=0):
        self.L = load_library()
        h = ctypes.c_void_p()
        check(self.L.coffe_engine_create(device, ctypes.byref(h)), "coffe_engine_create")
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            self.L.coffe_engine_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(self.L.coffe_engine_synchronize(self.h), "synchronize")


class Batch:
    """A fixed batch layout: ragged spectra, tables, separations, (z, l) rows."""

    def __init__(self, engine, n_in, n_fft, kmin, kmax, tables, sep, nz, nl):
        self.engine = engine
        L = engine.L
        self.nspec = len(n_in)
        self.n_in = np.asarray(n_in, dtype=np.int64)
        self.n_fft = np.asarray(n_fft, dtype=np.int64)
        self.kmin = np.ascontiguousarray(kmin, dtype=np.float64)
        self.kmax = np.ascontiguousarray(kmax, dtype=np.float64)
        self.tables = list(tables)
        self.sep = np.ascontiguousarray(sep, dtype=np.float64)
        self.nz, self.nl = nz, nl
        tn = (ctypes.c_int * len(self.tables))(*[t[0] for t in self.tables])
        tl = (ctypes.c_int * len(self.tables))(*[t[1] for t in self.tables])
        h = ctypes.c_void_p()
        dp = ctypes.POINTER(ctypes.c_double)
        i64p = ctypes.POINTER(ctypes.c_int64)
        check(L.coffe_batch_create(engine.h, self.nspec, self.n_in.ctypes.data_as(i64p),
                                   self.n_fft.ctypes.data_as(i64p), self.kmin.ctypes.data_as(dp),
                                   self.kmax.ctypes.data_as(dp), len(self.tables), tn, tl, len(self.sep),
                                   self.sep.ctypes.data_as(dp), nz, nl, ctypes.byref(h)), "coffe_batch_create")
        self.h = h
        self.L = L
        self.offsets = np.concatenate([[0], np.cumsum(self.n_in)])
        self.n_total = int(self.offsets[-1])

    def close(self):
        if self.h:
            self.L.coffe_batch_destroy(self.h)
            self.h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def _check_io(self, k, pk, A, itab, out):
        ntab, nsep = len(self.tables), len(self.sep)
        _need(k, (self.n_total,), "k")
        _need(pk, (self.n_total,), "pk")
        _need(A, (self.nz, self.nl, ntab, nsep), "A")
        _need(itab, (self.nspec, ntab, nsep), "itab")
        _need(out, (self.nspec, self.nz, self.nl, nsep), "out")

    @staticmethod
    def _stream():
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def forward(self, k, pk, A=None, itab=None, out=None, stream=None):
        """k, pk: float64 cuda tensors (concatenated); A: (nz, nl, ntab, nsep)."""
        import torch
        ntab, nsep = len(self.tables), len(self.sep)
        if itab is None:
            itab = torch.empty((self.nspec, ntab, nsep), dtype=torch.float64, device=k.device)
        if A is not None and out is None:
            out = torch.empty((self.nspec, self.nz, self.nl, nsep), dtype=torch.float64, device=k.device)
        self._check_io(k, pk, A, itab, out)
        check(self.L.coffe_batch_forward(self.h, _ptr(k), _ptr(pk), _ptr(A), _ptr(itab), _ptr(out),
                                         stream if stream is not None else self._stream()), "coffe_batch_forward")
        return itab, out

    def vjp(self, k, pk, A, gout, itab=None, out=None, gpk=None, stream=None):
        import torch
        ntab, nsep = len(self.tables), len(self.sep)
        if itab is None:
            itab = torch.empty((self.nspec, ntab, nsep), dtype=torch.float64, device=k.device)
        if out is None:
            out = torch.empty((self.nspec, self.nz, self.nl, nsep), dtype=torch.float64, device=k.device)
        if gpk is None:
            gpk = torch.empty_like(pk)
        self._check_io(k, pk, A, itab, out)
        _need(gout, (self.nspec, self.nz, self.nl, nsep), "gout")
        _need(gpk, (self.n_total,), "gpk")
        check(self.L.coffe_batch_vjp(self.h, _ptr(k), _ptr(pk), _ptr(A), _ptr(gout), _ptr(itab), _ptr(out),
                                     _ptr(gpk), stream if stream is not None else self._stream()), "coffe_batch_vjp")
        return itab, out, gpk

    # stage-by-stage launches (same kernels as forward / vjp), for pipelining and timing
    def stage_prepare(self, k, pk, stream):
        self._check_io(k, pk, None, None, None)
        check(self.L.coffe_batch_prepare(self.h, _ptr(k), _ptr(pk), stream), "coffe_batch_prepare")

    def stage_tables(self, A, gout, itab, stream):
        ntab, nsep = len(self.tables), len(self.sep)
        self._check_io(None, None, A, itab, None)
        _need(gout, (self.nspec, self.nz, self.nl, nsep), "gout")
        check(self.L.coffe_batch_tables_stage(self.h, _ptr(A), _ptr(gout), _ptr(itab), stream),
              "coffe_batch_tables_stage")

    def stage_project(self, A, itab, out, stream):
        self._check_io(None, None, A, itab, out)
        check(self.L.coffe_batch_project(self.h, _ptr(A), _ptr(itab), _ptr(out), stream), "coffe_batch_project")

    def stage_prepare_adj(self, k, pk, gpk, stream):
        self._check_io(k, pk, None, None, None)
        _need(gpk, (self.n_total,), "gpk")
        check(self.L.coffe_batch_prepare_adj(self.h, _ptr(k), _ptr(pk), _ptr(gpk), stream),
              "coffe_batch_prepare_adj")

    # the single kernels behind the prepare stages
    def kernel_resample(self, k, pk, stream):
        self._check_io(k, pk, None, None, None)
        check(self.L.coffe_batch_resample(self.h, _ptr(k), _ptr(pk), stream), "coffe_batch_resample")

    def kernel_fft_prepare(self, stream):
        check(self.L.coffe_batch_fft_prepare(self.h, stream), "coffe_batch_fft_prepare")

    def kernel_adj_c2r(self, stream):
        check(self.L.coffe_batch_adj_c2r(self.h, stream), "coffe_batch_adj_c2r")

    def kernel_spline_adj(self, k, gpk, stream):
        self._check_io(k, None, None, None, None)
        _need(gpk, (self.n_total,), "gpk")
        check(self.L.coffe_batch_spline_adj(self.h, _ptr(k), _ptr(gpk), stream), "coffe_batch_spline_adj")

    def small_r(self, k, pk, out=None, stream=None):
        """The 43 small-separation nodes [nspec][ntab][43]: r = 0 and the 42 min_sep values
        (integrals.c:520-557).  Synchronises the stream."""
        import torch
        if out is None:
            out = torch.empty((self.nspec, len(self.tables), 43), dtype=torch.float64, device=k.device)
        self._check_io(k, pk, None, None, None)
        _need(out, (self.nspec, len(self.tables), 43), "out")
        check(self.L.coffe_batch_small_r(self.h, _ptr(k), _ptr(pk), _ptr(out),
                                         stream if stream is not None else self._stream()), "coffe_batch_small_r")
        return out

    def tables_full(self, k, pk, stream=None):
        import torch
        ntab = len(self.tables)
        full = torch.zeros((self.nspec, ntab, int(self.n_fft.max())), dtype=torch.float64, device=k.device)
        self._check_io(k, pk, None, None, None)
        check(self.L.coffe_batch_tables(self.h, _ptr(k), _ptr(pk), _ptr(full),
                                        stream if stream is not None else self._stream()), "coffe_batch_tables")
        return full


def twofast_1bessel(k, pk, N, l, nu, r0=1.0, k0=None, kmin=None, kmax=None):
    """Drop-in twofast_1bessel (twofast.c:362-483) through the C-ABI; host arrays."""
    L = load_library()
    k = np.ascontiguousarray(k, dtype=np.float64)
    pk = np.ascontiguousarray(pk, dtype=np.float64)
    kmin = float(k[0]) if kmin is None else kmin
    kmax = float(k[-1]) if kmax is None else kmax
    k0 = kmin if k0 is None else k0
    ox, oy = np.empty(N), np.empty(N)
    dp = ctypes.POINTER(ctypes.c_double)
    check(L.twofast_1bessel_status(ox.ctypes.data_as(dp), oy.ctypes.data_as(dp), N, k.ctypes.data_as(dp),
                                   pk.ctypes.data_as(dp), len(k), float(l), float(nu), r0, k0, kmin, kmax),
          "twofast_1bessel")
    return ox, oy


def integral_table(n, l, k, pk, N, kmin=None, kmax=None):
    """The reference's full (N + 43)-point table (coffe_integrals_renormalizable,
    integrals.c:459-573) through coffe_integral_table; host arrays."""
    L = load_library()
    k = np.ascontiguousarray(k, dtype=np.float64)
    pk = np.ascontiguousarray(pk, dtype=np.float64)
    kmin = float(k[0]) if kmin is None else kmin
    kmax = float(k[-1]) if kmax is None else kmax
    ox, oy = np.empty(N + 43), np.empty(N + 43)
    dp = ctypes.POINTER(ctypes.c_double)
    check(L.coffe_integral_table(int(n), int(l), int(N), k.ctypes.data_as(dp), pk.ctypes.data_as(dp), len(k), kmin,
                                 kmax, ox.ctypes.data_as(dp), oy.ctypes.data_as(dp)), "coffe_integral_table")
    return ox, oy


def integral_table_state(n, l, state_n, state_l, k, pk, N, kmin=None, kmax=None):
    """The full (N + 43)-point table with the reference's integer states
    (coffe_integral_table_state; state 1 = half-integer: the table of (n/2, l/2)); host
    arrays."""
    L = load_library()
    k = np.ascontiguousarray(k, dtype=np.float64)
    pk = np.ascontiguousarray(pk, dtype=np.float64)
    kmin = float(k[0]) if kmin is None else kmin
    kmax = float(k[-1]) if kmax is None else kmax
    ox, oy = np.empty(N + 43), np.empty(N + 43)
    dp = ctypes.POINTER(ctypes.c_double)
    check(L.coffe_integral_table_state(int(n), int(l), int(state_n), int(state_l), int(N), k.ctypes.data_as(dp),
                                       pk.ctypes.data_as(dp), len(k), kmin, kmax, ox.ctypes.data_as(dp),
                                       oy.ctypes.data_as(dp)), "coffe_integral_table_state")
    return ox, oy


def half_integer_table(n2, l2, k, pk, N, kmin=None, kmax=None):
    """The half-integer table (n, l) = (n2/2, l2/2) of the flat-sky lensing terms
    (integrals.c:917-960, 1246-1300 with 459-573): r = 0, the 42 small separations and
    the N FFTLog nodes, r^(n - l) I_l^n(r) for n >= l."""
    return integral_table_state(n2, l2, 1, 1, k, pk, N, kmin, kmax)


FL_MAXK = 5   # half-integer tables per multipole of the density-lensing term (l <= 8)


def flat_lensing(zpar, ls, sep, bg, ll_tables, dl_tables, Omega_m, H0h, want_ll, want_dl, panels=128):
    """coffe_flat_lensing_multipoles: the flat-sky lensing-lensing and density-lensing
    multipoles (each times 2 pi^2, as signal.c adds them) summed into out[nz][nl][nsep].
    zpar[nz][7] = z_mean, chi, D1, s1, s2, b1, b2 at z_mean; bg the background dict (z,
    chi, D1); ll_tables {l: (x, y)} the (1, l) tables; dl_tables {(l, k): (x, y)} the
    half-integer tables ((l - 2k + 3)/2, (l - 2k + 1)/2).  The x integral takes `panels`
    panels of 32 Gauss-Legendre nodes: the Akima lookups make the integrand piecewise
    cubic with a kink at every table node below r, and 128 panels reach ~1e-10 of the
    integral up to r ~ 300 Mpc (16 panels: ~5e-9), where the reference's QAG stops
    at 1e-5."""
    L = load_library()
    zpar = np.ascontiguousarray(zpar, dtype=np.float64).reshape(-1, 7)
    ls_a = np.ascontiguousarray(ls, dtype=np.int32)
    sep = np.ascontiguousarray(sep, dtype=np.float64)
    nz, nl, nsep = len(zpar), len(ls_a), len(sep)
    xs, ys = [], []
    off = 0
    ll_off = np.zeros(nl, dtype=np.int64)
    ll_len = np.zeros(nl, dtype=np.int32)
    dl_off = np.zeros(nl * FL_MAXK, dtype=np.int64)
    dl_len = np.zeros(nl * FL_MAXK, dtype=np.int32)

    def add(tab):
        nonlocal off
        x, y = (np.ascontiguousarray(v, dtype=np.float64) for v in tab)
        xs.append(x)
        ys.append(y)
        start = off
        off += len(x)
        return start, len(x)

    for il, l in enumerate(ls_a):
        l = int(l)
        if l in ll_tables:
            ll_off[il], ll_len[il] = add(ll_tables[l])
        for k in range(FL_MAXK):
            if (l, k) in dl_tables:
                dl_off[il * FL_MAXK + k], dl_len[il * FL_MAXK + k] = add(dl_tables[(l, k)])
    tx = np.concatenate(xs) if xs else np.zeros(1)
    ty = np.concatenate(ys) if ys else np.zeros(1)
    bz, bc, bd = (np.ascontiguousarray(bg[key], dtype=np.float64) for key in ("z", "chi", "D1"))
    out = np.empty((nz, nl, nsep))
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    i64p = ctypes.POINTER(ctypes.c_int64)
    check(L.coffe_flat_lensing_multipoles(nz, zpar.ctypes.data_as(dp), nl, ls_a.ctypes.data_as(ip), nsep,
                                          sep.ctypes.data_as(dp), len(bz), bz.ctypes.data_as(dp),
                                          bc.ctypes.data_as(dp), bd.ctypes.data_as(dp), off, tx.ctypes.data_as(dp),
                                          ty.ctypes.data_as(dp), ll_off.ctypes.data_as(i64p),
                                          ll_len.ctypes.data_as(ip), dl_off.ctypes.data_as(i64p),
                                          dl_len.ctypes.data_as(ip), float(Omega_m), float(H0h), int(bool(want_ll)),
                                          int(bool(want_dl)), int(panels), out.ctypes.data_as(dp)),
          "coffe_flat_lensing_multipoles")
    return out


def flat_lensing_corrfunc(zpar, mus, sep, bg, table, Omega_m, H0h, want_ll, want_dl, panels=128):
    """coffe_flat_lensing_corrfunc: the flat-sky lensing-lensing and density-lensing
    correlation function out[nz][nmu][nsep]; table = the (1/2, -1/2) half-integer table
    (half_integer_table(1, -1, ...)); zpar as for flat_lensing."""
    L = load_library()
    zpar = np.ascontiguousarray(zpar, dtype=np.float64).reshape(-1, 7)
    mus = np.ascontiguousarray(mus, dtype=np.float64)
    sep = np.ascontiguousarray(sep, dtype=np.float64)
    tx, ty = (np.ascontiguousarray(v, dtype=np.float64) for v in table)
    bz, bc, bd = (np.ascontiguousarray(bg[key], dtype=np.float64) for key in ("z", "chi", "D1"))
    out = np.empty((len(zpar), len(mus), len(sep)))
    dp = ctypes.POINTER(ctypes.c_double)
    check(L.coffe_flat_lensing_corrfunc(len(zpar), zpar.ctypes.data_as(dp), len(mus), mus.ctypes.data_as(dp), len(sep),
                                        sep.ctypes.data_as(dp), len(bz), bz.ctypes.data_as(dp), bc.ctypes.data_as(dp),
                                        bd.ctypes.data_as(dp), len(tx), tx.ctypes.data_as(dp), ty.ctypes.data_as(dp),
                                        float(Omega_m), float(H0h), int(bool(want_ll)), int(bool(want_dl)),
                                        int(panels), out.ctypes.data_as(dp)), "coffe_flat_lensing_corrfunc")
    return out


# the contributions of coffe_integrated_corrfunc (COFFE_INTEGRATED_*)
INTEGRATED_FLAGS = {"den": 1, "rsd": 2, "d1": 4, "d2": 8, "g1": 16, "g2": 32, "g3": 64, "len": 128, "g4": 256,
                    "g5": 512, "only_cross": 1024, "flat_ll": 2048, "flat_dl": 4096}


def integrated_corrfunc(bg, bias_z, bias1, bias2, mb_z, s1, s2, fevo1, fevo2, Omega_m, H0h, tables, flags, zmean,
                        mus, sep, ren=None, ren0=None, rule=None):
    """coffe_integrated_corrfunc: the single- and double-integrated correlation function,
    full sky, at every (z_mean, mu, sep); returns (single, double), each (nz, nmu, nsep).
    bg: dict of host arrays z, chi, D1, a, f, cH, G1, G2; tables: {(n, l): (x, y)} (the
    standard tables and (4, 0), the divergent r^4 I_0^4); flags: names of INTEGRATED_FLAGS;
    ren: (chi, grid) the renormalisation grid and ren0: (sep, values) the r = 0
    renormalisation (with g4 / g5); rule: None or (p1, n1, pu, nu, ps, ns)."""
    L = load_library()
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    keep = []

    def arr(a):
        v = f64(a)
        keep.append(v)
        return v.ctypes.data_as(dp), len(v)

    unknown = set(flags) - set(INTEGRATED_FLAGS)
    if unknown:
        raise ValueError(f"unknown contributions {sorted(unknown)}")
    mask = 0
    for name in flags:
        mask |= INTEGRATED_FLAGS[name]
    inp = MuInputs()
    for name, key in (("bg_z", "z"), ("bg_chi", "chi"), ("bg_D1", "D1"), ("bg_f", "f"), ("bg_cH", "cH"),
                      ("bg_G1", "G1"), ("bg_G2", "G2"), ("bg_a", "a")):
        ptr, n = arr(bg[key])
        setattr(inp, name, ptr)
    inp.nbg = n
    inp.bias_z, inp.nb = arr(bias_z)
    inp.bias1, _ = arr(bias1)
    inp.bias2, _ = arr(bias2)
    inp.mb_z, inp.nmb = arr(mb_z)
    for name, v in (("s1", s1), ("s2", s2), ("fevo1", fevo1), ("fevo2", fevo2)):
        setattr(inp, name, arr(v)[0])
    inp.H0h, inp.Omega_m = float(H0h), float(Omega_m)
    n0, x0, y0 = 0, None, None
    if ren is not None:
        inp.ren_chi, inp.nren = arr(ren[0])
        inp.ren, _ = arr(np.asarray(ren[1]).reshape(-1))
    if ren0 is not None:
        x0, n0 = arr(ren0[0])
        y0, _ = arr(ren0[1])
    keys = list(tables)
    tn = (ctypes.c_int * len(keys))(*[int(t[0]) for t in keys])
    tl = (ctypes.c_int * len(keys))(*[int(t[1]) for t in keys])
    tlen = (ctypes.c_int * len(keys))(*[len(tables[t][0]) for t in keys])
    tx, _ = arr(np.concatenate([f64(tables[t][0]) for t in keys]) if keys else np.zeros(1))
    ty, _ = arr(np.concatenate([f64(tables[t][1]) for t in keys]) if keys else np.zeros(1))
    zm, mu, sp = f64(np.atleast_1d(zmean)), f64(np.atleast_1d(mus)), f64(np.atleast_1d(sep))
    r = (ctypes.c_int * 6)(*rule) if rule is not None else None
    out = np.empty((len(zm), len(mu), len(sp), 2))
    check(L.coffe_integrated_corrfunc(ctypes.byref(inp), ctypes.c_uint(mask), len(keys), tn, tl, tlen, tx, ty, n0, x0,
                                      y0, len(zm), zm.ctypes.data_as(dp), len(mu), mu.ctypes.data_as(dp), len(sp),
                                      sp.ctypes.data_as(dp), r, out.ctypes.data_as(dp)), "coffe_integrated_corrfunc")
    return out[..., 0], out[..., 1]


# separations of the divergent table I_0^4, in Mpc (integrals.c:76-99)
COFFE_SEP = np.array(
    [0.000001, 0.000005, 0.0001, 0.005, 0.01, 0.02, 0.05, 0.1, 0.2, 0.4, 0.8, 1., 1.5]
    + list(range(2, 31)) + list(range(32, 81, 2)) + list(range(81, 96))
    + [95.5, 96., 96.5, 97., 97.5, 98., 98.5, 99., 99.5, 100., 100.5, 101., 101.5, 102., 102.5, 103., 103.5,
       104., 104.5, 105.] + list(range(106, 111)) + list(range(112, 121, 2)) + list(range(124, 181, 4))
    + list(range(185, 241, 5)) + list(range(250, 301, 10)) + list(range(320, 701, 20))
    + list(range(750, 6601, 50)) + list(range(7000, 20001, 1000)), dtype=np.float64)


def divergent_tables(k, pk, sep=None, chi=None, kmin=None, kmax=None, return_status=False):
    """The divergent table I_0^4 and its renormalisations (integrals.c:772-909) through
    coffe_divergent_tables, host arrays: returns (sep, I_0^4(sep) sep^4, the r = 0
    renormalisation at sep times sep^4, the chi x chi grid ren[a, b] = (chi_b, chi_a)),
    all divided by 2 pi^2.  sep defaults to the reference's separation list, chi to no
    grid.  return_status adds the per-integral QUADPACK codes and interval counts (in
    the order sep for I_0^4, sep for the renormalisation, then the grid) and returns
    the values also when an integral reached the interval cap."""
    L = load_library()
    k = np.ascontiguousarray(k, dtype=np.float64)
    pk = np.ascontiguousarray(pk, dtype=np.float64)
    sep = COFFE_SEP.copy() if sep is None else np.ascontiguousarray(sep, dtype=np.float64)
    chi = np.zeros(0) if chi is None else np.ascontiguousarray(chi, dtype=np.float64)
    kmin = float(k[0]) if kmin is None else kmin
    kmax = float(k[-1]) if kmax is None else kmax
    i40, r0 = np.empty(len(sep)), np.empty(len(sep))
    ren = np.empty((len(chi), len(chi)))
    units = 2 * len(sep) + len(chi) ** 2
    status = np.empty(units, dtype=np.int32)
    nint = np.empty(units, dtype=np.int32)
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    rc = L.coffe_divergent_tables(k.ctypes.data_as(dp), pk.ctypes.data_as(dp), len(k), kmin, kmax,
                                  sep.ctypes.data_as(dp), len(sep), chi.ctypes.data_as(dp), len(chi),
                                  i40.ctypes.data_as(dp), r0.ctypes.data_as(dp), ren.ctypes.data_as(dp),
                                  status.ctypes.data_as(ip), nint.ctypes.data_as(ip))
    if return_status and rc == COFFE_ENOCONV:
        return sep, i40, r0, ren, status, nint
    check(rc, "coffe_divergent_tables")
    if return_status:
        return sep, i40, r0, ren, status, nint
    return sep, i40, r0, ren


def mu_weights_jvp(engine, bg, tangent, bias_z, bias1, bias2, zmean, ls, tables, sep, den=True, rsd=True, d1=False,
                   flatsky=False, nmu=256, terms=None, stream=None):
    """(A, At): the weights of mu_weights (density / RSD / Doppler-1 terms) and their
    tangents along one direction of the background (coffe_mu_weights_jvp).  tangent: the
    derivatives of the tables chi, D1, f, cH, G1, G2 at the nodes bg["z"] (from
    coffe_amd.background.jvp); the biases have none.  Both (nz, nl, ntab, nsep) device
    tensors; nl <= 4."""
    import torch
    L = engine.L
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
    dp = ctypes.POINTER(ctypes.c_double)
    keep = []

    def arr(a):
        v = f64(a)
        keep.append(v)
        return v.ctypes.data_as(dp), len(v)

    inp, tan = MuInputs(), MuInputs()
    for name, key in (("bg_z", "z"), ("bg_chi", "chi"), ("bg_D1", "D1"), ("bg_f", "f"), ("bg_cH", "cH"),
                      ("bg_G1", "G1"), ("bg_G2", "G2")):
        ptr, n = arr(bg[key])
        setattr(inp, name, ptr)
        if key != "z":
            tp, tn = arr(tangent[key])
            if tn != n:
                raise ValueError(f"the tangent of {key} must have one value per node")
            setattr(tan, name, tp)
    inp.nbg = tan.nbg = n
    inp.bias_z, inp.nb = arr(bias_z)
    inp.bias1, _ = arr(bias1)
    inp.bias2, _ = arr(bias2)
    zm, sp = f64(np.atleast_1d(zmean)), f64(np.atleast_1d(sep))
    ls_c = (ctypes.c_int * len(ls))(*ls)
    tn_c = (ctypes.c_int * len(tables))(*[t[0] for t in tables])
    tl_c = (ctypes.c_int * len(tables))(*[t[1] for t in tables])
    if terms is None:
        terms = terms_of(den, rsd, d1)
    shape = (len(zm), len(ls), len(tables), len(sp))
    A = torch.empty(shape, dtype=torch.float64, device="cuda")
    At = torch.empty(shape, dtype=torch.float64, device="cuda")
    st = stream if stream is not None else ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(L.coffe_mu_weights_jvp(engine.h, ctypes.byref(inp), ctypes.byref(tan), int(terms), int(flatsky), 0,
                                 len(zm), zm.ctypes.data_as(dp), len(ls), ls_c, len(tables), tn_c, tl_c, len(sp),
                                 sp.ctypes.data_as(dp), int(nmu), _ptr(A), _ptr(At), st), "coffe_mu_weights_jvp")
    return A, At


def akima(x, y, xq):
    """Host Akima interpolation through the engine library (coffe_akima_eval)."""
    L = load_library()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    xq = np.ascontiguousarray(np.atleast_1d(xq), dtype=np.float64)
    out = np.empty(len(xq))
    dp = ctypes.POINTER(ctypes.c_double)
    check(L.coffe_akima_eval(len(x), x.ctypes.data_as(dp), y.ctypes.data_as(dp), len(xq), xq.ctypes.data_as(dp),
                             out.ctypes.data_as(dp)), "coffe_akima_eval")
    return out


TERM_DEN_DEN, TERM_RSD_RSD, TERM_D1_D1, TERM_DEN_RSD, TERM_DEN_D1, TERM_RSD_D1 = 1, 2, 4, 8, 16, 32
TERM_DG_DG, TERM_DEN_DG, TERM_RSD_DG, TERM_D1_DG = 64, 128, 256, 512
TERM_DG_MUTUAL = 1024   # with TERM_DG_DG: the mutual d2 / g terms only, their autos left out


def terms_of(den, rsd, d1, dg=False):
    """COFFE_TERM_* mask of the contribution flags (auto terms and cross terms); dg: any of
    d2, g1-g3 (their mutual terms and their crosses with den, rsd, d1)."""
    t = 0
    t |= TERM_DEN_DEN if den else 0
    t |= TERM_RSD_RSD if rsd else 0
    t |= TERM_D1_D1 if d1 else 0
    t |= TERM_DEN_RSD if den and rsd else 0
    t |= TERM_DEN_D1 if den and d1 else 0
    t |= TERM_RSD_D1 if rsd and d1 else 0
    if dg:
        t |= TERM_DG_DG
        t |= TERM_DEN_DG if den else 0
        t |= TERM_RSD_DG if rsd else 0
        t |= TERM_D1_DG if d1 else 0
    return t


_DP = ctypes.POINTER(ctypes.c_double)


class MuInputs(ctypes.Structure):
    """coffe_mu_inputs (include/coffe_engine.h)."""
    _fields_ = [("nbg", ctypes.c_int)] + [(n, _DP) for n in ("bg_z", "bg_chi", "bg_D1", "bg_f", "bg_cH", "bg_G1",
                                                              "bg_G2", "bg_a")] + \
        [("nb", ctypes.c_int)] + [(n, _DP) for n in ("bias_z", "bias1", "bias2")] + \
        [("nmb", ctypes.c_int)] + [(n, _DP) for n in ("mb_z", "s1", "s2", "fevo1", "fevo2")] + \
        [("H0h", ctypes.c_double), ("Omega_m", ctypes.c_double), ("nren", ctypes.c_int), ("ren_chi", _DP),
         ("ren", _DP)]


def mu_weights(engine, bg, bias_z, bias1, bias2, zmean, ls, tables, sep, den=True, rsd=True, d1=False,
               flatsky=False, nmu=256, terms=None, out=None, stream=None, dg=None, ren_out=None, paired=False,
               rule=None):
    """Spectrum-independent mu-quadrature weights A[z][l][t][r] (coffe_mu_weights_ex), so that
    multipoles = sum_t A I_t(sep) (+ the renormalisation part with the d2 / g terms).
    bg: dict of host arrays z, chi, D1, f, cH (conformal Hubble rate), G1, G2 (and a with
    dg).  terms: COFFE_TERM_* mask, default from den/rsd/d1 and dg.
    dg: None or a dict with sources (a set of "d2", "g1", "g2", "g3"), H0h, Omega_m, the
    magnification / evolution biases mb_z, s1, s2, fevo1, fevo2, and for the auto terms the
    renormalisation grid ren_chi, ren (I_0^4's table (4, 0) must be among tables); then
    ren_out (device, (nz, nl, nsep), or (nz, nl) paired) receives the grid's part.
    paired: point i is (zmean[i], sep[i]); A is (nz, nl, ntab).
    rule: None (the nmu-node Gauss-Legendre rule) or (mu nodes, weights) through
    coffe_mu_weights_rule; one node mu with weight 2 and l = 0 gives the correlation
    function at mu."""
    import torch
    L = engine.L
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
    dp = ctypes.POINTER(ctypes.c_double)
    keep = []

    def arr(a):
        v = f64(a)
        keep.append(v)
        return v.ctypes.data_as(dp), len(v)

    inp = MuInputs()
    for name, key in (("bg_z", "z"), ("bg_chi", "chi"), ("bg_D1", "D1"), ("bg_f", "f"), ("bg_cH", "cH"),
                      ("bg_G1", "G1"), ("bg_G2", "G2")):
        ptr, n = arr(bg[key])
        setattr(inp, name, ptr)
    inp.nbg = n
    inp.bias_z, inp.nb = arr(bias_z)
    inp.bias1, _ = arr(bias1)
    inp.bias2, _ = arr(bias2)
    sources = 0
    if dg is not None:
        src = dg["sources"]
        sources = (1 if "d2" in src else 0) | (2 if "g1" in src else 0) | (4 if "g2" in src else 0) | \
            (8 if "g3" in src else 0)
        inp.bg_a, _ = arr(bg["a"])
        inp.mb_z, inp.nmb = arr(dg["mb_z"])
        for name in ("s1", "s2", "fevo1", "fevo2"):
            setattr(inp, name, arr(dg[name])[0])
        inp.H0h, inp.Omega_m = float(dg["H0h"]), float(dg["Omega_m"])
        if "ren" in dg:
            inp.ren_chi, inp.nren = arr(dg["ren_chi"])
            inp.ren, _ = arr(np.asarray(dg["ren"]).reshape(-1))
    zm, sp = f64(np.atleast_1d(zmean)), f64(np.atleast_1d(sep))
    if paired and len(zm) != len(sp):
        raise ValueError("paired weights need as many redshifts as separations")
    ls_c = (ctypes.c_int * len(ls))(*ls)
    tn = (ctypes.c_int * len(tables))(*[t[0] for t in tables])
    tl = (ctypes.c_int * len(tables))(*[t[1] for t in tables])
    if terms is None:
        terms = terms_of(den, rsd, d1, dg is not None)
    shape = (len(zm), len(ls), len(tables)) if paired else (len(zm), len(ls), len(tables), len(sp))
    A = out if out is not None else torch.empty(shape, dtype=torch.float64, device="cuda")
    _need(A, shape, "A")
    if ren_out is not None:
        _need(ren_out, shape[:2] if paired else (len(zm), len(ls), len(sp)), "ren_out")
    st = stream if stream is not None else ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    ren_p = _ptr(ren_out) if ren_out is not None else None
    if rule is not None:
        rmu, rw = f64(np.atleast_1d(rule[0])), f64(np.atleast_1d(rule[1]))
        if len(rmu) != len(rw):
            raise ValueError("the rule needs one weight per node")
        check(L.coffe_mu_weights_rule(engine.h, ctypes.byref(inp), int(terms), int(sources), int(flatsky),
                                      int(paired), len(zm), zm.ctypes.data_as(dp), len(ls), ls_c, len(tables), tn,
                                      tl, len(sp), sp.ctypes.data_as(dp), len(rmu), rmu.ctypes.data_as(dp),
                                      rw.ctypes.data_as(dp), _ptr(A), ren_p, st), "coffe_mu_weights_rule")
        return A
    check(L.coffe_mu_weights_ex(engine.h, ctypes.byref(inp), int(terms), int(sources), int(flatsky), int(paired),
                                len(zm), zm.ctypes.data_as(dp), len(ls), ls_c, len(tables), tn, tl, len(sp),
                                sp.ctypes.data_as(dp), int(nmu), _ptr(A), ren_p, st), "coffe_mu_weights")
    return A
