"""A stand-in for the COFFE C program around the reference entry points of
include/coffe_refabi.h: ctypes mirrors of its structures, the host's interpolation
routines (coffe_init_spline and friends, common.c) registered through
coffe_refabi_set_spline_ops, and a builder that fills coffe_parameters_t /
coffe_background_t from a configured coffe_amd.coffe.Coffe object.

The host's splines are GSL-like objects (gsl_spline: interp, x, y, size; gsl_spline2d:
the gsl_interp2d head, then xarr, yarr, zarr) whose node arrays the library reads, and
interp_spline evaluates them with the engine's host Akima (GSL's akima, the reference's
default interp_method) or a natural cubic spline (GSL's cspline)."""
import ctypes

import numpy as np

D, I, SZ, VP = ctypes.c_double, ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p
DP, IP = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
STR = ctypes.c_char * 256

COFFE_H0 = 1 / 2997.92458
CORRFUNC, MULTIPOLES, AVERAGE_MULTIPOLES, COVARIANCE_MULTIPOLES, COVARIANCE_AVERAGE_MULTIPOLES = 1, 2, 3, 4, 5
NONINTEGRATED, SINGLE_INTEGRATED, DOUBLE_INTEGRATED = 0, 1, 2
INTERP_CSPLINE, INTERP_AKIMA = 3, 5


class Interp(ctypes.Structure):
    _fields_ = [("spline", VP), ("accel", VP)]


class Interp2d(ctypes.Structure):
    _fields_ = [("spline", VP), ("xaccel", VP), ("yaccel", VP)]


class Contrib(ctypes.Structure):
    _fields_ = [(n, I) for n in ("den", "rsd", "len", "d1", "d2", "g1", "g2", "g3", "g4", "g5")]


class FitArray(ctypes.Structure):
    _fields_ = [("array", VP), ("size", SZ)]


class Parameters(ctypes.Structure):
    """coffe_parameters_t of the C build (common.h:340-550)."""
    _fields_ = (
        [("output_type", I), ("mu", DP), ("mu_len", SZ), ("correlation_contrib", Contrib),
         ("type_bg", ctypes.POINTER(ctypes.c_char_p)), ("type_bg_len", SZ), ("background_bins", I),
         ("bessel_bins", I)]
        + [(n, D) for n in ("inv_k_window", "Omega0_cdm", "Omega0_m")]
        + [(n, I) for n in ("b_derivative", "f_derivative", "b_tilde_derivative", "f_tilde_derivative")]
        + [(n, D) for n in ("Omega0_baryon", "Omega0_gamma", "w0", "wa", "Omega0_de")]
        + [("z_mean", DP), ("z_mean_len", SZ), ("deltaz", DP), ("deltaz_len", SZ), ("sep", DP), ("sep_len", SZ)]
        + [(n, I) for n in ("integration_method", "integration_bins", "nthreads")]
        + [("file_power_spectrum", STR), ("power_spectrum", Interp), ("power_spectrum_norm", Interp),
           ("power_spectrum2d", Interp2d), ("power_spectrum2d_norm", Interp2d)]
        + [(n, D) for n in ("k_min", "k_max", "k_min_norm", "k_max_norm")]
        + [("read_galaxy_bias1", I), ("read_galaxy_bias2", I), ("file_galaxy_bias1", STR), ("file_galaxy_bias2", STR)]
        + [(f"galaxy_bias{i}", Interp) for i in (1, 2, 3, 4)]
        + [("read_magnification_bias1", I), ("read_magnification_bias2", I), ("galaxy_bias_analytic", I),
           ("file_magnification_bias1", STR), ("file_magnification_bias2", STR),
           ("magnification_bias1", Interp), ("magnification_bias2", Interp),
           ("read_evolution_bias1", I), ("read_evolution_bias2", I),
           ("file_evolution_bias1", STR), ("file_evolution_bias2", STR),
           ("evolution_bias1", Interp), ("evolution_bias2", Interp), ("divergent", I), ("conf", VP),
           ("timestamp", STR), ("output_path", STR), ("output_prefix", STR), ("interp_method", I),
           ("multipole_values", IP), ("multipole_values_len", SZ), ("has_class", I), ("has_cuba", I),
           ("zmin", DP), ("zmin_len", SZ), ("zmax", DP), ("zmax_len", SZ), ("fsky", DP), ("fsky_len", SZ),
           ("density1", DP), ("density2", DP), ("density1_len", SZ), ("density2_len", SZ),
           ("pixelsize", DP), ("pixelsize_len", SZ)]
        + [(f"covariance_pop{i}", I) for i in (1, 2, 3, 4)]
        + [(n, I) for n in ("covariance_cosmic", "covariance_mixed", "covariance_poisson",
                            "covariance_integration_method", "covariance_integration_bins",
                            "covariance_interpolation_method", "covariance_window", "have_window")]
        + [("window_size", D)]
        + [(n, I) for n in ("flatsky_local", "flatsky_local_nonlocal", "flatsky_nonlocal", "verbose", "have_class")]
        + [(n, D) for n in ("n_s", "sigma8", "A_s", "T_cmb", "N_ur", "m_ncdm", "Omega0_nu", "YHe")]
        + [("N_ncdm", I), ("h", D), ("k_pivot", D), ("class_struct", VP * 8),
           ("class_precision", type("ClassPrecision", (ctypes.Structure,), {"_fields_": [
               ("k_per_decade_for_bao", D), ("k_per_decade_for_pk", D), ("start_large_k_at_tau_h_over_tau_k", D),
               ("l_max_g", I), ("l_max_ur", I), ("tol_perturb_integration", D),
               ("radiation_streaming_trigger_tau_over_tau_k", D), ("ur_fluid_trigger_tau_over_tau_k", D)]})),
           ("flag", I), ("pk_type", I), ("midpoint_approximation", I), ("only_cross_correlations", I),
           ("integration_1d_type", I), ("integration_1d_prec", D)]
        + [(n + "_coefficients", FitArray) for n in ("galaxy_bias1", "galaxy_bias2", "magnification_bias1",
                                                     "magnification_bias2", "evolution_bias1", "evolution_bias2")]
        + [("degree_" + n, I) for n in ("galaxy_bias1", "galaxy_bias2", "magnification_bias1",
                                         "magnification_bias2", "evolution_bias1", "evolution_bias2")]
    )


class Background(ctypes.Structure):
    """coffe_background_t (background.h:23-44)."""
    _fields_ = [(n, Interp) for n in ("z_as_chi", "a", "Hz", "conformal_Hz", "conformal_Hz_prime", "D1", "f", "G1",
                                      "G2", "comoving_distance")] + [("flag", I)]


class Integral(ctypes.Structure):
    _fields_ = [("result", Interp), ("result2d", Interp2d), ("renormalization", Interp2d),
                ("renormalization_zero_separation", Interp), ("n", I), ("l", I), ("state_n", I), ("state_l", I)]


class IntegralArray(ctypes.Structure):
    _fields_ = [("array", ctypes.POINTER(Integral)), ("size", SZ)]


class MultipolesCoords(ctypes.Structure):
    _fields_ = [("z_mean", D), ("separation", D), ("l", I)]


class Multipoles(ctypes.Structure):
    _fields_ = [("coords", MultipolesCoords), ("value", D)]


class MultipolesArray(ctypes.Structure):
    _fields_ = [("array", ctypes.POINTER(Multipoles)), ("size", SZ)]


class AverageMultipolesCoords(ctypes.Structure):
    _fields_ = [("z_min", D), ("z_max", D), ("separation", D), ("l", I)]


class AverageMultipoles(ctypes.Structure):
    _fields_ = [("coords", AverageMultipolesCoords), ("value", D)]


class AverageMultipolesArray(ctypes.Structure):
    _fields_ = [("array", ctypes.POINTER(AverageMultipoles)), ("size", SZ)]


class CovarianceCoords(ctypes.Structure):
    _fields_ = [("l1", I), ("l2", I), ("separation1", D), ("separation2", D), ("z_mean", D)]


class Covariance(ctypes.Structure):
    _fields_ = [("coords", CovarianceCoords), ("value", D)]


class CovarianceArray(ctypes.Structure):
    _fields_ = [("array", ctypes.POINTER(Covariance)), ("size", SZ)]


class AverageCovarianceCoords(ctypes.Structure):
    _fields_ = [("l1", I), ("l2", I), ("separation1", D), ("separation2", D), ("z_min", D), ("z_max", D)]


class AverageCovariance(ctypes.Structure):
    _fields_ = [("coords", AverageCovarianceCoords), ("value", D)]


class AverageCovarianceArray(ctypes.Structure):
    _fields_ = [("array", ctypes.POINTER(AverageCovariance)), ("size", SZ)]


class GslSpline(ctypes.Structure):
    _fields_ = [("interp", VP), ("x", DP), ("y", DP), ("size", SZ)]


class GslSpline2d(ctypes.Structure):
    _fields_ = [("type", VP), ("xmin", D), ("xmax", D), ("ymin", D), ("ymax", D), ("xsize", SZ), ("ysize", SZ),
                ("state", VP), ("xarr", DP), ("yarr", DP), ("zarr", DP)]


INIT = ctypes.CFUNCTYPE(I, ctypes.POINTER(Interp), DP, DP, SZ, I)
INIT2D = ctypes.CFUNCTYPE(I, ctypes.POINTER(Interp2d), DP, DP, DP, SZ, SZ, I)
EVAL = ctypes.CFUNCTYPE(D, ctypes.POINTER(Interp), D)
EVAL2D = ctypes.CFUNCTYPE(D, ctypes.POINTER(Interp2d), D, D)
FREE = ctypes.CFUNCTYPE(I, ctypes.POINTER(Interp))
FREE2D = ctypes.CFUNCTYPE(I, ctypes.POINTER(Interp2d))


class SplineOps(ctypes.Structure):
    _fields_ = [("init_spline", INIT), ("init_spline2d", INIT2D), ("interp_spline", EVAL),
                ("interp_spline2d", EVAL2D), ("free_spline", FREE), ("free_spline2d", FREE2D)]


class Host:
    """The host program's splines: every object lives in self.objects (keyed by its
    address) with its node arrays, until freed."""

    def __init__(self, lib):
        self.lib = lib
        self.objects = {}
        self.freed = 0
        self.ops = SplineOps(INIT(self._init), INIT2D(self._init2d), EVAL(self._eval), EVAL2D(self._eval2d),
                             FREE(self._free), FREE2D(self._free2d))
        lib.coffe_refabi_set_spline_ops.argtypes = [ctypes.POINTER(SplineOps)]
        lib.coffe_refabi_set_spline_ops(ctypes.byref(self.ops))

    def close(self):
        self.lib.coffe_refabi_set_spline_ops(None)

    # -- the routines ---------------------------------------------------------------
    def _new(self, x, y, kind):
        x = np.array(x, dtype=np.float64)
        y = np.array(y, dtype=np.float64)
        s = GslSpline(None, x.ctypes.data_as(DP), y.ctypes.data_as(DP), len(x))
        addr = ctypes.addressof(s)
        self.objects[addr] = (s, x, y, kind)
        return addr

    def spline(self, x, y, kind=INTERP_AKIMA):
        """A coffe_interpolation of the host (as coffe_init_spline would make it)."""
        return Interp(self._new(x, y, kind), None)

    def _init(self, interp, x, y, n, kind):
        interp.contents.spline = self._new(np.ctypeslib.as_array(x, (n,)), np.ctypeslib.as_array(y, (n,)), kind)
        interp.contents.accel = None
        return 0

    def _init2d(self, interp, x, y, z, nx, ny, kind):
        xa = np.array(np.ctypeslib.as_array(x, (nx,)))
        ya = np.array(np.ctypeslib.as_array(y, (ny,)))
        za = np.array(np.ctypeslib.as_array(z, (nx * ny,)))
        s = GslSpline2d(None, xa[0], xa[-1], ya[0], ya[-1], nx, ny, None, xa.ctypes.data_as(DP),
                        ya.ctypes.data_as(DP), za.ctypes.data_as(DP))
        addr = ctypes.addressof(s)
        self.objects[addr] = (s, xa, ya, za, kind)
        interp.contents.spline = addr
        return 0

    def _eval(self, interp, v):
        from coffe_amd.engine import akima
        s, x, y, kind = self.objects[interp.contents.spline]
        if kind == INTERP_AKIMA:
            return float(akima(x, y, [v])[0])
        from scipy.interpolate import CubicSpline
        return float(CubicSpline(x, y, bc_type="natural")(v))

    def _eval2d(self, interp, a, b):
        return float("nan")   # the library reads the grid's arrays, never evaluates it

    def _free(self, interp):
        if interp.contents.spline:
            self.objects.pop(interp.contents.spline, None)
            self.freed += 1
        interp.contents.spline = None
        return 0

    def _free2d(self, interp):
        if interp.contents.spline:
            self.objects.pop(interp.contents.spline, None)
            self.freed += 1
        interp.contents.spline = None
        return 0

    def nodes(self, interp):
        """(x, y) node arrays of a host spline."""
        obj = self.objects[interp.spline]
        return obj[1], obj[2]

    def grid(self, interp2d):
        obj = self.objects[interp2d.spline]
        return obj[1], obj[2], obj[3]


def declare(lib):
    P = ctypes.POINTER
    lib.coffe_integrals_init.argtypes = [P(Parameters), P(Background), P(IntegralArray)]
    lib.coffe_integrals_free.argtypes = [P(IntegralArray)]
    lib.coffe_find_integral.argtypes = [P(IntegralArray), I, I, I, I]
    lib.coffe_find_integral.restype = P(Integral)
    lib.coffe_integrate.argtypes = [P(Parameters), P(Background), P(IntegralArray), D, D, D, D, D, I, I, I]
    lib.coffe_integrate.restype = D
    lib.coffe_multipoles_init.argtypes = [P(Parameters), P(Background), P(IntegralArray), P(MultipolesArray)]
    lib.coffe_multipoles_free.argtypes = [P(MultipolesArray)]
    lib.coffe_average_multipoles_init.argtypes = [P(Parameters), P(Background), P(IntegralArray),
                                                  P(AverageMultipolesArray)]
    lib.coffe_average_multipoles_free.argtypes = [P(AverageMultipolesArray)]
    lib.coffe_covariance_init.argtypes = [P(Parameters), P(Background), P(CovarianceArray),
                                          P(AverageCovarianceArray)]
    lib.coffe_covariance_free.argtypes = [P(CovarianceArray)]
    lib.coffe_average_covariance_free.argtypes = [P(AverageCovarianceArray)]
    lib.coffe_covariance_find.argtypes = [P(CovarianceArray), D, I, I, D, D]
    lib.coffe_covariance_find.restype = Covariance
    lib.covariance_coefficient.argtypes = [D, D, D, I]
    lib.covariance_coefficient.restype = D
    return lib


class Program:
    """coffe_parameters_t and coffe_background_t of a configured Coffe object, as the
    reference's parser and background module would fill them (the arrays are owned
    here)."""

    def __init__(self, host, cosmo, output_type=MULTIPOLES, bessel_bins=None, divergent=None):
        self.host, self.keep = host, []
        p = cosmo._p
        par = Parameters()
        par.output_type = output_type
        c = par.correlation_contrib
        c.den, c.rsd, c.d1 = int(p["has_density"]), int(p["has_rsd"]), int(p["has_d1"])
        c.d2, c.g1, c.g2, c.g3 = (int(p[k]) for k in ("has_d2", "has_g1", "has_g2", "has_g3"))
        c.len = int(p["has_lensing"])
        c.g4, c.g5 = int(p["has_g4"]), int(p["has_g5"])
        par.bessel_bins = int(p["integrals_fftlog_sampling"] if bessel_bins is None else bessel_bins)
        par.Omega0_cdm, par.Omega0_baryon, par.Omega0_gamma = p["omega_cdm"], p["omega_baryon"], p["omega_gamma"]
        par.Omega0_m = p["omega_cdm"] + p["omega_baryon"]
        par.w0, par.wa, par.h = p["w0"], p["wa"], p["h"]
        par.z_mean, par.z_mean_len = self._arr(p["z_mean"])
        par.deltaz, par.deltaz_len = self._arr(p["deltaz"])
        par.sep, par.sep_len = self._arr(p["sep"])
        ls = np.ascontiguousarray(p["l"], dtype=np.int32)
        self.keep.append(ls)
        par.multipole_values, par.multipole_values_len = ls.ctypes.data_as(IP), len(ls)
        k, pk = cosmo._pk
        par.power_spectrum = host.spline(k, pk)
        par.power_spectrum_norm = host.spline(k, pk)
        par.k_min = par.k_min_norm = k[0]
        par.k_max = par.k_max_norm = k[-1]
        for i in (1, 2, 3, 4):
            b = cosmo._bias[i]
            setattr(par, f"galaxy_bias{i}", host.spline(b.x, b.y))
        for name, key in (("magnification_bias1", "s1"), ("magnification_bias2", "s2"),
                          ("evolution_bias1", "fevo1"), ("evolution_bias2", "fevo2")):
            b = cosmo._mbias[key]
            setattr(par, name, host.spline(b.x, b.y))
        dg = any(p[k] for k in ("has_d2", "has_g1", "has_g2", "has_g3", "has_g4", "has_g5"))
        par.divergent = int(dg if divergent is None else divergent)
        par.interp_method = INTERP_AKIMA
        par.zmin, par.zmin_len = self._arr(p["z_min"])
        par.zmax, par.zmax_len = self._arr(p["z_max"])
        par.fsky, par.fsky_len = self._arr(p["fsky"])
        par.density1, par.density1_len = self._arr(p["number_density1"])
        par.density2, par.density2_len = self._arr(p["number_density2"])
        par.pixelsize, par.pixelsize_len = self._arr(p["pixelsize"])
        pops = cosmo.covariance_populations
        par.covariance_pop1, par.covariance_pop2, par.covariance_pop3, par.covariance_pop4 = (int(v) for v in pops)
        par.covariance_cosmic = int(p["covariance_cosmic"])
        par.covariance_mixed = int(p["covariance_mixed"])
        par.covariance_poisson = int(p["covariance_poisson"])
        par.covariance_integration_method = 2
        par.covariance_integration_bins = int(p["covariance_integration_sampling"])
        par.covariance_interpolation_method = 2
        par.covariance_window = int(p["has_binned_covariance"])
        par.flatsky_local = int(p["has_flatsky_local"])
        par.flatsky_nonlocal = int(p["has_flatsky_nonlocal"])
        par.flatsky_local_nonlocal = int(p["has_flatsky_local_nonlocal"])
        par.only_cross_correlations = int(p["has_only_cross_correlations"])
        par.mu, par.mu_len = self._arr(p["mu"])
        par.integration_1d_type, par.integration_1d_prec = 1, 1e-5
        self.par = par
        bg = cosmo._background()
        B = Background()
        for name, key in (("a", "a"), ("Hz", "Hz"), ("conformal_Hz", "conformal_Hz"),
                          ("conformal_Hz_prime", "conformal_Hz_prime"), ("D1", "D1"), ("f", "f"), ("G1", "G1"),
                          ("G2", "G2"), ("comoving_distance", "chi")):
            setattr(B, name, host.spline(bg["z"], bg[key]))
        B.z_as_chi = host.spline(bg["chi"], bg["z"])
        self.bg = B

    def _arr(self, v):
        a = np.ascontiguousarray(np.atleast_1d(np.asarray(v, dtype=np.float64)))
        self.keep.append(a)
        return a.ctypes.data_as(DP), len(a)
