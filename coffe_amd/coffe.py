"""Drop-in for COFFE's Python API (python/coffe/coffe.pyx) on the spherical-Bessel
path: the Coffe object with its parameters, settings files (from_file / to_file),
set_power_spectrum_linear, biases, background lookups, integral(r, n, l), the
correlation function, the multipoles, their redshift averages and the Gaussian
covariance of both, computed by the MI355X engine.

Units follow the reference: k in 1/Mpc, P in Mpc^3, separations in Mpc.  Parameters
are validated as the reference validates them (TypeError / ValueError), lazily
computed pieces are dropped when a parameter they depend on changes
(coffe.pyx:948-993), and results come back as Multipoles / Covariance / ... records in
the reference's order (multipoles.c:76-86: z, l, r; covariance.c:1046-1060: z, l1, r1,
l2, r2).  The terms are density, RSD, Doppler-1, d2 and g1-g3 with their cross terms, and the
integrated ones: lensing, g4 and g5 with each other and with the local terms (full
sky, or the flat-sky lensing forms with has_flatsky_nonlocal /
has_flatsky_local_nonlocal).  Massive neutrinos and the CLASS / nonlinear spectra are
not part of this engine and raise NotImplementedError; the
parameters of the CLASS interface and of the reference's CPU integrators are kept and
written to settings files but change nothing here.
"""
import warnings
from configparser import ConfigParser
from pathlib import Path

import numpy as np

from . import background as _background
from .representation import AverageCovariance, AverageMultipoles, Corrfunc, Covariance, Multipoles

TABLES = [(0, 0), (0, 2), (0, 4), (1, 1), (1, 3), (2, 0), (2, 2), (3, 1)]   # integrals.c:599-608
_DG = ("d2", "g1", "g2", "g3")
COFFE_H0 = 1 / 2997.92458   # h/Mpc (common.h)

# the named choices of the enum-valued parameters (coffe.pyx:213-233, common.h:283-310)
_CHOICES = {
    "pk_type": {0: "linear", 1: "linear_class", 2: "halofit", 3: "hmcode"},
    "covariance_integration_method": {1: "standard", 2: "fftlog"},
    "covariance_interpolation_method": {1: "bilinear", 2: "bicubic"},
    "integration_1d_type": {1: "gsl", 2: "exponential"},
}

_MAX_SIZE = 2 ** 63 - 1
# Parameters of the reference's CLASS interface and of its CPU integrators
# (coffe.pyx:1306-1545, 2439-2520): checked as the reference checks them, kept, listed
# by `parameters` and written by to_file, with no effect on what this engine computes.
_KEPT = {
    "n_s": ((int, float), -100, 100), "sigma8": ((int, float), -100, 100), "A_s": ((int, float), -100, 100),
    "T_cmb": ((int, float), -100, 100), "N_ur": ((int, float), -100, 100), "YHe": ((int, float), -100, 100),
    "N_ncdm": (int, -100, 100), "integration_1d_prec": (float, 0, 1), "integration_sampling": (int, 0, _MAX_SIZE),
    "k_per_decade_for_bao": ((int, float), None, None), "k_per_decade_for_pk": ((int, float), None, None),
    "start_large_k_at_tau_h_over_tau_k": ((int, float), None, None), "l_max_g": (int, None, None),
    "l_max_ur": (int, None, None), "tol_perturb_integration": ((int, float), None, None),
    "radiation_streaming_trigger_tau_over_tau_k": ((int, float), None, None),
    "ur_fluid_trigger_tau_over_tau_k": ((int, float), None, None),
}


class CoffeConfigParser(ConfigParser):
    """The settings-file parser (coffe.pyx:59-82): arrays are written "[a, b, ...]"."""

    def getfloat_array(self, section, option, *args, **kwargs):
        return [float(v) for v in self.get(section, option).strip("[]").split(",")]

    def getint_array(self, section, option, *args, **kwargs):
        return [int(v) for v in self.get(section, option).strip("[]").split(",")]

    def getstr_array(self, section, option, *args, **kwargs):
        return [str(v.strip(' "')) for v in self.get(section, option).strip("[]").split(",")]


# Options of older settings files and what replaces them (coffe.pyx:571-589); None:
# read and ignored, with a DeprecationWarning either way.
_LEGACY = [
    ("only_cross_correlations", "int", "has_only_cross_correlations"),
    ("covariance_z_mean", "float_array", "z_mean"),
    ("covariance_deltaz", "float_array", "deltaz"),
    ("covariance_fsky", "float_array", "fsky"),
    ("covariance_pixelsize", "float_array", "pixelsize"),
    ("covariance_step_size", "float", None),
    ("covariance_zmin", "float_array", "z_min"),
    ("covariance_zmax", "float_array", "z_max"),
    ("covariance_minimum_separation", "float", None),
    ("covariance_window", "int", "has_binned_covariance"),
    ("covariance_density", "float_array", ["number_density1", "number_density2"]),
    ("covariance_integration_bins", "int", None),
    ("multipoles", "int_array", "l"),
    ("flatsky_local", "int", "has_flatsky_local"),
    ("flatsky_local_nonlocal", "int", "has_flatsky_local_nonlocal"),
    ("flatsky_nonlocal", "int", "has_flatsky_nonlocal"),
]
_CONTRIBUTION_NAMES = {"den": "has_density", "rsd": "has_rsd", "len": "has_lensing", "d1": "has_d1",
                       "d2": "has_d2", "g1": "has_g1", "g2": "has_g2", "g3": "has_g3", "g4": "has_g4",
                       "g5": "has_g5"}
_FILE_FLOAT_ARRAYS = ("sep", "z_mean", "deltaz", "z_min", "z_max", "mu", "number_density1", "number_density2",
                      "pixelsize", "fsky")
_FILE_INT_ARRAYS = ("l", "covariance_populations")
_FILE_FLOATS = ("omega_m", "omega_baryon", "omega_gamma", "w0", "wa", "h", "sigma8", "n_s", "T_cmb", "A_s",
                "k_per_decade_for_bao", "k_per_decade_for_pk", "start_large_k_at_tau_h_over_tau_k",
                "tol_perturb_integration", "radiation_streaming_trigger_tau_over_tau_k",
                "ur_fluid_trigger_tau_over_tau_k", "N_ur", "m_ncdm", "k_min", "k_max", "YHe", "integration_1d_prec")
_FILE_INTS = ("has_density", "has_lensing", "has_rsd", "has_d1", "has_d2", "has_g1", "has_g2", "has_g3", "has_g4",
              "has_g5", "integration_sampling", "background_sampling", "has_binned_covariance", "has_flatsky_local",
              "has_flatsky_nonlocal", "has_flatsky_local_nonlocal", "has_only_cross_correlations",
              "covariance_cosmic", "covariance_mixed", "covariance_poisson", "covariance_integration_sampling",
              "integrals_fftlog_sampling", "N_ncdm", "l_max_g", "l_max_ur", "multipole_quadrature_nodes", "verbose")
_FILE_INT_OR_STR = ("covariance_integration_method", "pk_type", "covariance_interpolation_method",
                    "integration_1d_type")


def _check_parameter(name, value, kind, xmin=None, xmax=None):
    """Type and range check with the reference's exceptions (coffe.pyx:86-110)."""
    if not isinstance(value, kind) or isinstance(value, bool):
        raise TypeError(f"Expected {kind} for {name}, got {type(value)}")
    if xmin is not None and value < xmin:
        raise ValueError(f"Value {value} of {name} smaller than xmin ({xmin})")
    if xmax is not None and value > xmax:
        raise ValueError(f"Value {value} of {name} larger than xmax ({xmax})")


def _float_list(name, value, ok):
    try:
        iter(value)
    except TypeError as err:
        raise TypeError(f"The value {value} of {name} is not iterable") from err
    try:
        out = [float(v) for v in value]
    except (TypeError, ValueError) as err:
        raise TypeError(f"Cannot convert all values of {name} to floats") from err
    if not all(ok(v) for v in out):
        raise ValueError(f"{name}: values out of range")
    return np.array(out)


class _Spline:
    """A sampled function evaluated by Akima (the default interp_method, parser.c:807)."""

    def __init__(self, x, y):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)

    def __call__(self, v):
        from .engine import akima
        return akima(self.x, self.y, v)


class Coffe:
    """Engine-backed COFFE object; the keyword arguments go to set_parameters."""

    _defaults = dict(
        omega_cdm=0.25, omega_baryon=0.05, omega_gamma=9e-5, h=0.67, w0=-1.0, wa=0.0,
        sep=5.0 + 3.0 * np.arange(100), l=[0, 2, 4], z_mean=[1.0], deltaz=[0.2], fsky=[1.0], pixelsize=[10.0],
        z_min=[0.9], z_max=[1.1], mu=[0.5],
        number_density1=[1e-3], number_density2=[1e-3],
        has_density=True, has_rsd=False, has_d1=False, has_d2=False, has_g1=False, has_g2=False, has_g3=False, has_g4=False, has_g5=False,
        has_flatsky_local=False, has_only_cross_correlations=False,
        has_lensing=False, has_flatsky_nonlocal=False, has_flatsky_local_nonlocal=False,
        integrals_fftlog_sampling=10000, covariance_integration_sampling=8000,
        covariance_cosmic=True, covariance_mixed=True, covariance_poisson=True,
        has_binned_covariance=False, covariance_populations=[1, 1, 1, 1],
        multipole_quadrature_nodes=256,
        # the window of P(k) the integrals use (parser.c:840-841); set_power_spectrum_linear
        # moves it to the ends of the samples (coffe.pyx:2974-2977)
        k_min=1e-5, k_max=300.0, background_sampling=10000, verbose=False,
        pk_type=0, covariance_integration_method=1, covariance_interpolation_method=2, integration_1d_type=1,
        # parser.c:750-799, 994: values of the CLASS interface and the CPU integrators
        integration_1d_prec=1e-5, integration_sampling=750000,
        n_s=0.96, sigma8=0.8156, A_s=0.0, T_cmb=2.726, N_ur=2.0328, YHe=0.0, N_ncdm=1, m_ncdm=0.0,
        k_per_decade_for_bao=70.0, k_per_decade_for_pk=10.0, start_large_k_at_tau_h_over_tau_k=0.07,
        l_max_g=12, l_max_ur=17, tol_perturb_integration=1e-5,
        radiation_streaming_trigger_tau_over_tau_k=45.0, ur_fluid_trigger_tau_over_tau_k=30.0,
    )

    def __init__(self, **kwargs):
        self._p = {k: (np.array(v) if isinstance(v, (list, np.ndarray)) else v) for k, v in self._defaults.items()}
        zb = np.linspace(0.0, 15.0, 16)
        self._bias = {i: _Spline(zb, np.ones_like(zb)) for i in (1, 2, 3, 4)}
        # magnification and evolution biases (parser.c:900-940: 0 by default)
        self._mbias = {k: _Spline(zb, np.zeros_like(zb)) for k in ("s1", "s2", "fevo1", "fevo2")}
        self._pk = None
        self._engine = None
        self._free_all()
        if kwargs:
            self.set_parameters(**kwargs)

    # ---- lazily computed state -------------------------------------------------
    def _free_all(self):
        self._bg = None
        self._lens = None
        self._lens_cf = None
        self._tables = None
        self._divergent = None
        self._divergent_r0 = None
        self._ren = None
        self._free_results()

    def _free_spectrum(self):
        """Drops what depends on P(k) or its window."""
        self._lens = None
        self._lens_cf = None
        self._tables = None
        self._divergent = None
        self._divergent_r0 = None
        self._ren = None
        self._free_results()

    def _free_results(self):
        self._multipoles = None
        self._average_multipoles = None
        self._covariance = None
        self._average_covariance = None
        self._corrfunc = None

    def _engine_handle(self):
        if self._engine is None:
            from .engine import Engine
            self._engine = Engine(0)
        return self._engine

    # ---- parameters --------------------------------------------------------------
    @property
    def parameters(self):
        """The settable parameters and their values, as set_parameters takes them
        (coffe.pyx:1136-1154)."""
        out = {}
        for k in list(self._p) + ["omega_m"]:
            v = getattr(self, k)
            out[k] = list(v) if isinstance(v, np.ndarray) else v
        return out

    def set_parameters(self, **kwargs):
        for k, v in kwargs.items():
            if not hasattr(type(self), k) or not isinstance(getattr(type(self), k), property):
                raise AttributeError(f"unknown parameter {k}")
            setattr(self, k, v)

    def _set_omegas(self, name, value, others, cdm, gamma):
        """The reference's bookkeeping (coffe.pyx:1178-1256): Omega_m = Omega_cdm +
        Omega_baryon; setting omega_baryon or omega_m keeps the other of the two and moves
        Omega_cdm; Omega_de closes the budget and must stay >= 0 for the values checked."""
        _check_parameter(name, value, (int, float), -100, 100)
        if np.allclose(value, getattr(self, name)):
            return
        if sum([float(value)] + list(others)) > 1:
            raise ValueError("Cannot satisfy the condition Omega_de >= 0")
        self._p["omega_cdm"], self._p["omega_gamma"] = float(cdm), float(gamma)
        if name == "omega_baryon":
            self._p["omega_baryon"] = float(value)
        self._free_all()

    omega_cdm = property(
        lambda s: s._p["omega_cdm"],
        lambda s, v: s._set_omegas("omega_cdm", v, (s.omega_baryon, s.omega_gamma), v, s.omega_gamma))
    omega_baryon = property(
        lambda s: s._p["omega_baryon"],
        lambda s, v: s._set_omegas("omega_baryon", v, (s.omega_cdm, s.omega_gamma), s.omega_m - v, s.omega_gamma))
    omega_gamma = property(
        lambda s: s._p["omega_gamma"],
        lambda s, v: s._set_omegas("omega_gamma", v, (s.omega_m,), s.omega_cdm, v))
    omega_m = property(
        lambda s: s._p["omega_cdm"] + s._p["omega_baryon"],
        lambda s, v: s._set_omegas("omega_m", v, (s.omega_gamma,), v - s.omega_baryon, s.omega_gamma))

    @property
    def omega_de(self):
        return 1 - self.omega_m - self._p["omega_gamma"]

    # only big Omegas here (use_little_omega is False in the reference, coffe.pyx:1014-1053)
    use_little_omega = property(lambda s: False)
    big_omega_m = property(lambda s: s.omega_m)
    big_omega_de = property(lambda s: s.omega_de)
    big_omega_baryon = property(lambda s: s.omega_baryon)
    big_omega_cdm = property(lambda s: s.omega_cdm)
    big_omega_gamma = property(lambda s: s.omega_gamma)
    big_omega_nu = property(lambda s: 0.0)
    omega_nu = property(lambda s: 0.0)
    # no CLASS and no Cuba here (coffe.pyx:1023-1029)
    has_class = property(lambda s: False)
    has_cuba = property(lambda s: False)

    def _set_scalar(self, name, value, kind, lo, hi, frees):
        _check_parameter(name, value, kind, lo, hi)
        if self._p[name] != value:
            self._p[name] = value
            frees()

    h = property(lambda s: s._p["h"], lambda s, v: s._set_scalar("h", v, (int, float), 1e-10, 100, s._free_all))
    w0 = property(lambda s: s._p["w0"], lambda s, v: s._set_scalar("w0", v, (int, float), -100, 100, s._free_all))
    wa = property(lambda s: s._p["wa"], lambda s, v: s._set_scalar("wa", v, (int, float), -100, 100, s._free_all))
    integrals_fftlog_sampling = property(
        lambda s: s._p["integrals_fftlog_sampling"],
        lambda s, v: s._set_scalar("integrals_fftlog_sampling", v, int, 4, 1000000, s._free_all))
    covariance_integration_sampling = property(
        lambda s: s._p["covariance_integration_sampling"],
        lambda s, v: s._set_scalar("covariance_integration_sampling", v, int, 8, 1000000, s._free_results))
    multipole_quadrature_nodes = property(
        lambda s: s._p["multipole_quadrature_nodes"],
        lambda s, v: s._set_scalar("multipole_quadrature_nodes", v, int, 2, 65536, s._free_results))

    def _set_list(self, name, value, ok, frees, as_int=False):
        if as_int:
            try:
                iter(value)
            except TypeError as err:
                raise TypeError(f"The value {value} of {name} is not iterable") from err
            try:
                arr = np.array([v.__index__() for v in value], dtype=int)
            except AttributeError as err:
                raise TypeError(f"Cannot convert all values of {name} to int") from err
            if not all(ok(v) for v in arr):
                raise ValueError(f"{name}: values out of range")
        else:
            arr = _float_list(name, value, ok)
        self._p[name] = arr
        frees()

    sep = property(lambda s: s._p["sep"].copy(),
                   lambda s, v: s._set_list("sep", v, lambda x: 0 < x < 10000, s._free_results))
    z_mean = property(lambda s: s._p["z_mean"].copy(),
                      lambda s, v: s._set_list("z_mean", v, lambda x: 0 <= x < 10, s._free_results))
    z_min = property(lambda s: s._p["z_min"].copy(),
                     lambda s, v: s._set_list("z_min", v, lambda x: 0 <= x < 10, s._free_results))
    z_max = property(lambda s: s._p["z_max"].copy(),
                     lambda s, v: s._set_list("z_max", v, lambda x: 0 <= x < 10, s._free_results))
    # the angles of the correlation function, in (-1, 1) (coffe.pyx:1589-1627)
    mu = property(lambda s: s._p["mu"].copy(),
                  lambda s, v: s._set_list("mu", v, lambda x: -1 < x < 1, s._free_results))
    l = property(lambda s: s._p["l"].copy(),  # noqa: E741
                 lambda s, v: s._set_list("l", v, lambda x: 0 <= x < 10, s._free_results, as_int=True))
    deltaz = property(lambda s: s._p["deltaz"].copy(),
                      lambda s, v: s._set_list("deltaz", v, lambda x: 0 < x < 10, s._free_results))
    fsky = property(lambda s: s._p["fsky"].copy(),
                    lambda s, v: s._set_list("fsky", v, lambda x: 0 < x <= 1, s._free_results))
    pixelsize = property(lambda s: s._p["pixelsize"].copy(),
                         lambda s, v: s._set_list("pixelsize", v, lambda x: 0 < x < 10000, s._free_results))
    # coffe.pyx:2097-2157: values in [0, 10); a zero density fails at compute time here
    number_density1 = property(lambda s: s._p["number_density1"].copy(),
                               lambda s, v: s._set_list("number_density1", v, lambda x: 0 <= x < 10, s._free_results))
    number_density2 = property(lambda s: s._p["number_density2"].copy(),
                               lambda s, v: s._set_list("number_density2", v, lambda x: 0 <= x < 10, s._free_results))

    @property
    def covariance_populations(self):
        """The populations of the four fields of the covariance (coffe.pyx:2160-2192)."""
        return self._p["covariance_populations"].copy()

    @covariance_populations.setter
    def covariance_populations(self, value):
        try:
            iter(value)
        except TypeError as err:
            raise TypeError(f"The value {value} is not iterable") from err
        try:
            temp = [int(v) for v in value]
        except (TypeError, ValueError) as err:
            raise TypeError("Cannot convert all values to integers") from err
        if len(temp) != 4:
            raise ValueError("The array for setting `covariance_populations` must have length 4")
        if any(v not in (1, 2, 3, 4) for v in value):
            raise ValueError("The values for `covariance_populations` must be one of: [1, 2, 3, 4]")
        self._p["covariance_populations"] = np.array(temp)
        self._free_results()

    def _set_flag(self, name, value):
        value = bool(value)
        if self._p[name] != value:
            self._p[name] = value
            self._free_results()

    has_density = property(lambda s: s._p["has_density"], lambda s, v: s._set_flag("has_density", v))
    has_rsd = property(lambda s: s._p["has_rsd"], lambda s, v: s._set_flag("has_rsd", v))
    has_d1 = property(lambda s: s._p["has_d1"], lambda s, v: s._set_flag("has_d1", v))
    has_d2 = property(lambda s: s._p["has_d2"], lambda s, v: s._set_flag("has_d2", v))
    has_g1 = property(lambda s: s._p["has_g1"], lambda s, v: s._set_flag("has_g1", v))
    has_g2 = property(lambda s: s._p["has_g2"], lambda s, v: s._set_flag("has_g2", v))
    has_g3 = property(lambda s: s._p["has_g3"], lambda s, v: s._set_flag("has_g3", v))
    # the integrated potential terms g4 (functions.c:3837-3851) and g5 (3853-3874) with
    # their cross terms
    has_g4 = property(lambda s: s._p["has_g4"], lambda s, v: s._set_flag("has_g4", v))
    has_g5 = property(lambda s: s._p["has_g5"], lambda s, v: s._set_flag("has_g5", v))
    has_flatsky_local = property(lambda s: s._p["has_flatsky_local"],
                                 lambda s, v: s._set_flag("has_flatsky_local", v))
    # lensing (signal.c:556-771): full sky by the integrated-terms kernel, or the flat-sky
    # lensing-lensing term with has_flatsky_nonlocal and the flat-sky density-lensing
    # term with has_flatsky_local_nonlocal
    has_lensing = property(lambda s: s._p["has_lensing"], lambda s, v: s._set_flag("has_lensing", v))
    has_flatsky_nonlocal = property(lambda s: s._p["has_flatsky_nonlocal"],
                                    lambda s, v: s._set_flag("has_flatsky_nonlocal", v))
    has_flatsky_local_nonlocal = property(lambda s: s._p["has_flatsky_local_nonlocal"],
                                          lambda s, v: s._set_flag("has_flatsky_local_nonlocal", v))
    # only the cross terms of the contributions (coffe.pyx:1121-1133, functions.c:129-669)
    has_only_cross_correlations = property(lambda s: s._p["has_only_cross_correlations"],
                                           lambda s, v: s._set_flag("has_only_cross_correlations", v))
    covariance_cosmic = property(lambda s: s._p["covariance_cosmic"], lambda s, v: s._set_flag("covariance_cosmic", v))
    covariance_mixed = property(lambda s: s._p["covariance_mixed"], lambda s, v: s._set_flag("covariance_mixed", v))
    covariance_poisson = property(lambda s: s._p["covariance_poisson"],
                                  lambda s, v: s._set_flag("covariance_poisson", v))
    # the bin-averaged Poisson term (covariance_window, coffe.pyx:2525-2545)
    has_binned_covariance = property(lambda s: s._p["has_binned_covariance"],
                                     lambda s, v: s._set_flag("has_binned_covariance", v))
    verbose = property(lambda s: s._p["verbose"], lambda s, v: s._p.__setitem__("verbose", bool(v)))

    # the window of P(k) the integral tables use (coffe.pyx:2895-2926)
    k_min = property(lambda s: s._p["k_min"],
                     lambda s, v: s._set_scalar("k_min", v, (int, float), 0, 1e-3, s._free_spectrum))
    k_max = property(lambda s: s._p["k_max"],
                     lambda s, v: s._set_scalar("k_max", v, (int, float), 0.1, 1e3, s._free_spectrum))
    # the number of background samples on [0, 15] (coffe.pyx:2505-2518)
    background_sampling = property(
        lambda s: s._p["background_sampling"],
        lambda s, v: s._set_scalar("background_sampling", v, int, 0, _MAX_SIZE, s._free_all))

    def _set_choice(self, name, value):
        """An enum-valued parameter, set by code or by name (coffe.pyx:236-245)."""
        table = _CHOICES[name]
        if value not in list(table) + list(table.values()):
            raise ValueError(f"The value `{name}={value}` is not one of `{list(table) + list(table.values())}`")
        code = value if value in table else {v: c for c, v in table.items()}[value]
        if name == "pk_type" and code != 0:
            raise NotImplementedError("only the linear power spectrum (pk_type 'linear') is computed by this "
                                      "engine: the CLASS and nonlinear spectra are not part of it")
        self._p[name] = code
        self._free_results()

    # The reference chooses among its CPU methods with these (coffe.pyx:2242-2325,
    # 2438-2460).  The engine computes the covariance integrals by 2D FFTLog for either
    # method (the reference's tests hold its default, 'standard', to the same fixtures at
    # rtol 5e-4) and the table nodes by adaptive Gauss-Kronrod for either 1D type; the
    # covariance is interpolated bicubically only ('bilinear' raises at compute time).
    pk_type = property(lambda s: _CHOICES["pk_type"][s._p["pk_type"]],
                       lambda s, v: s._set_choice("pk_type", v))
    covariance_integration_method = property(
        lambda s: _CHOICES["covariance_integration_method"][s._p["covariance_integration_method"]],
        lambda s, v: s._set_choice("covariance_integration_method", v))
    covariance_interpolation_method = property(
        lambda s: _CHOICES["covariance_interpolation_method"][s._p["covariance_interpolation_method"]],
        lambda s, v: s._set_choice("covariance_interpolation_method", v))
    integration_1d_type = property(
        lambda s: _CHOICES["integration_1d_type"][s._p["integration_1d_type"]],
        lambda s, v: s._set_choice("integration_1d_type", v))

    @property
    def m_ncdm(self):
        """Mass of the non-CDM species in eV (coffe.pyx:1496-1513); only 0 here."""
        return self._p["m_ncdm"]

    @m_ncdm.setter
    def m_ncdm(self, value):
        _check_parameter("m_ncdm", value, (int, float), -100, 100)
        if value != 0:
            raise NotImplementedError("massive neutrinos (m_ncdm != 0) are not part of this engine")
        self._p["m_ncdm"] = float(value)

    def reset_contributions(self):
        """Sets every writable has_* parameter to False (coffe.pyx:2702-2708)."""
        for name in self._p:
            if name.startswith("has_"):
                setattr(self, name, False)

    # ---- inputs ----------------------------------------------------------------------
    def set_power_spectrum_linear(self, k, pk, z=0):
        """The linear P(k) at z = 0 (k in 1/Mpc, P in Mpc^3); sets k_min/k_max to its ends
        (coffe.pyx:2929-2996)."""
        k = np.asarray(k, dtype=np.float64)
        pk = np.asarray(pk, dtype=np.float64)
        if not np.allclose(k, np.sort(k)):
            raise ValueError("The input k array must be sorted")
        if len(k) != len(pk):
            raise ValueError("The input arrays have mismatching lengths")
        if len(k) < 5:
            raise ValueError("The power spectrum needs at least 5 samples")
        self._pk = (np.ascontiguousarray(k), np.ascontiguousarray(pk))
        self._p["k_min"], self._p["k_max"] = float(k[0]), float(k[-1])
        self._free_spectrum()

    def _kwin(self):
        """The window [k_min, k_max] of the integrals, inside the P(k) samples."""
        self._need_pk()
        k = self._pk[0]
        kmin, kmax = float(self._p["k_min"]), float(self._p["k_max"])
        if not (k[0] <= kmin < kmax <= k[-1]):
            raise ValueError(f"[k_min, k_max] = [{kmin}, {kmax}] is not inside the power spectrum's samples "
                             f"[{k[0]}, {k[-1]}]")
        return kmin, kmax

    def power_spectrum(self, k, z):
        """P(k, z) = D1(z)^2 P(k) (Akima of the input samples)."""
        self._need_pk()
        _check_parameter("k", k, (int, float), self.k_min, self.k_max)
        return float(_Spline(*self._pk)(float(k))[0] * self.growth_factor(z) ** 2)

    def cross_spectrum(self, k, z1, z2, approximation="geometric"):
        """P(k) D1(z1) D1(z2) (coffe.pyx:2822-2860)."""
        self._need_pk()
        _check_parameter("k", k, (int, float), self.k_min, self.k_max)
        _check_parameter("z1", z1, (int, float), 0, 15)
        _check_parameter("z2", z2, (int, float), 0, 15)
        return float(_Spline(*self._pk)(float(k))[0] * self.growth_factor(z1) * self.growth_factor(z2))

    def _set_bias(self, i, x, y):
        x = np.asarray(x, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        if not np.allclose(x, np.sort(x)):
            raise ValueError("The input array must be sorted")
        if len(x) != len(y):
            raise ValueError(f"Mismatching lengths for x ({len(x)}) and y ({len(y)})")
        if len(x) < 5:
            raise ValueError("At least 5 samples are needed")
        _check_parameter("xmin", float(x[0]), (int, float), xmin=0)
        _check_parameter("xmax", float(x[-1]), (int, float), xmax=15)
        self._bias[i] = _Spline(x, y)
        self._free_results()

    def _eval_bias(self, i, z):
        s = self._bias[i]
        _check_parameter("x", z, (int, float), float(s.x[0]), float(s.x[-1]))
        return float(s(float(z))[0])

    def set_galaxy_bias1(self, x_sampling, y_sampling):
        self._set_bias(1, x_sampling, y_sampling)

    def set_galaxy_bias2(self, x_sampling, y_sampling):
        self._set_bias(2, x_sampling, y_sampling)

    def set_galaxy_bias3(self, x_sampling, y_sampling):
        self._set_bias(3, x_sampling, y_sampling)

    def set_galaxy_bias4(self, x_sampling, y_sampling):
        self._set_bias(4, x_sampling, y_sampling)

    def galaxy_bias1(self, z):
        return self._eval_bias(1, z)

    def galaxy_bias2(self, z):
        return self._eval_bias(2, z)

    def galaxy_bias3(self, z):
        return self._eval_bias(3, z)

    def galaxy_bias4(self, z):
        return self._eval_bias(4, z)

    def _set_mbias(self, key, x, y):
        x = np.asarray(x, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        if not np.allclose(x, np.sort(x)):
            raise ValueError("The input array must be sorted")
        if len(x) != len(y):
            raise ValueError(f"Mismatching lengths for x ({len(x)}) and y ({len(y)})")
        if len(x) < 5:
            raise ValueError("At least 5 samples are needed")
        self._mbias[key] = _Spline(x, y)
        # the background's G1 / G2 read them (background.c:301-697)
        self._bg = None
        self._free_results()

    def _eval_mbias(self, key, z):
        s = self._mbias[key]
        _check_parameter("x", z, (int, float), float(s.x[0]), float(s.x[-1]))
        return float(s(float(z))[0])

    def set_magnification_bias1(self, x_sampling, y_sampling):
        """Magnification bias s of the first population (coffe.pyx:1837-1857)."""
        self._set_mbias("s1", x_sampling, y_sampling)

    def set_magnification_bias2(self, x_sampling, y_sampling):
        self._set_mbias("s2", x_sampling, y_sampling)

    def set_evolution_bias1(self, x_sampling, y_sampling):
        """Evolution bias f_evo of the first population (coffe.pyx:1892-1957)."""
        self._set_mbias("fevo1", x_sampling, y_sampling)

    def set_evolution_bias2(self, x_sampling, y_sampling):
        self._set_mbias("fevo2", x_sampling, y_sampling)

    def magnification_bias1(self, z):
        return self._eval_mbias("s1", z)

    def magnification_bias2(self, z):
        return self._eval_mbias("s2", z)

    def evolution_bias1(self, z):
        return self._eval_mbias("fevo1", z)

    def evolution_bias2(self, z):
        return self._eval_mbias("fevo2", z)

    # The sampled biases as to_file writes them: the spline's size evenly spaced points
    # over its range and the spline's values there (coffe.pyx:845-945).
    def _samples(self, s):
        return np.linspace(float(s.x[0]), float(s.x[-1]), len(s.x))

    def galaxy_bias1_redshifts(self):
        return self._samples(self._bias[1])

    def galaxy_bias1_values(self):
        return np.array([self.galaxy_bias1(float(z)) for z in self.galaxy_bias1_redshifts()])

    def galaxy_bias2_redshifts(self):
        return self._samples(self._bias[2])

    def galaxy_bias2_values(self):
        return np.array([self.galaxy_bias2(float(z)) for z in self.galaxy_bias2_redshifts()])

    def galaxy_bias3_redshifts(self):
        return self._samples(self._bias[3])

    def galaxy_bias3_values(self):
        return np.array([self.galaxy_bias3(float(z)) for z in self.galaxy_bias3_redshifts()])

    def galaxy_bias4_redshifts(self):
        return self._samples(self._bias[4])

    def galaxy_bias4_values(self):
        return np.array([self.galaxy_bias4(float(z)) for z in self.galaxy_bias4_redshifts()])

    def magnification_bias1_redshifts(self):
        return self._samples(self._mbias["s1"])

    def magnification_bias1_values(self):
        return np.array([self.magnification_bias1(float(z)) for z in self.magnification_bias1_redshifts()])

    def magnification_bias2_redshifts(self):
        return self._samples(self._mbias["s2"])

    def magnification_bias2_values(self):
        return np.array([self.magnification_bias2(float(z)) for z in self.magnification_bias2_redshifts()])

    def evolution_bias1_redshifts(self):
        return self._samples(self._mbias["fevo1"])

    def evolution_bias1_values(self):
        return np.array([self.evolution_bias1(float(z)) for z in self.evolution_bias1_redshifts()])

    def evolution_bias2_redshifts(self):
        return self._samples(self._mbias["fevo2"])

    def evolution_bias2_values(self):
        return np.array([self.evolution_bias2(float(z)) for z in self.evolution_bias2_redshifts()])

    # ---- settings files ---------------------------------------------------------------
    def to_file(self, filename):
        """Writes `key = value` for every parameter (booleans as 0 / 1, arrays as
        "[a, b, ...]") and then the sampled biases (coffe.pyx:794-843)."""
        lines = []
        for key, value in self.parameters.items():
            if isinstance(value, bool):
                value = int(value)
            elif isinstance(value, (list, tuple, np.ndarray)):
                value = [v.item() if isinstance(v, np.generic) else v for v in value]
            lines.append(f"{key} = {value}")
        for kind, pops in (("galaxy", (1, 2, 3, 4)), ("magnification", (1, 2)), ("evolution", (1, 2))):
            for i in pops:
                name = f"{kind}_bias{i}"
                lines.append(f"{name}_redshifts = {getattr(self, name + '_redshifts')().tolist()}")
                lines.append(f"{name}_values = {getattr(self, name + '_values')().tolist()}")
        Path(filename).write_text("\n".join(lines) + "\n", encoding="utf-8")

    @staticmethod
    def from_file(filename):
        """A Coffe object with the parameters of a settings file (coffe.pyx:536-791):
        the reference's `key = value;` files, including the older option names, which
        are mapped to the current ones with a DeprecationWarning.  The options of the
        reference's readers are read (also has_g2, has_only_cross_correlations,
        covariance_populations, l_max_g / l_max_ur as integers and this engine's
        multipole_quadrature_nodes); other keys are ignored."""
        section = "default"
        text = f"[{section}]\n" + "\n".join(line.strip(";") for line in Path(filename).read_text().split("\n"))
        parser = CoffeConfigParser()
        parser.read_string(text)
        config = parser[section]
        cosmo = Coffe()

        def read(kind, option):
            return getattr(config, f"get{kind}")(option)

        for name, kind, replaced_by in _LEGACY:
            if name not in config:
                continue
            if replaced_by is None:
                warnings.warn(f"Option '{name}' is no longer supported and will be ignored.", DeprecationWarning)
                continue
            warnings.warn(f"Option '{name}' is no longer supported; it may be replaced by '{replaced_by}' "
                          "(if found).", DeprecationWarning)
            for target in (replaced_by if isinstance(replaced_by, list) else [replaced_by]):
                setattr(cosmo, target, read(kind, name))
        if "correlation_contributions" in config:
            warnings.warn("Option 'correlation_contributions' is no longer supported; it may be replaced by one "
                          f"of {sorted(set(_CONTRIBUTION_NAMES.values()))} (if found).", DeprecationWarning)
            contributions = config.getstr_array("correlation_contributions")
            for key, target in _CONTRIBUTION_NAMES.items():
                if key in contributions:
                    setattr(cosmo, target, True)
        for options, kind in ((_FILE_FLOAT_ARRAYS, "float_array"), (_FILE_INT_ARRAYS, "int_array"),
                              (_FILE_FLOATS, "float"), (_FILE_INTS, "int")):
            for option in options:
                if option in config:
                    setattr(cosmo, option, read(kind, option))
        for option in _FILE_INT_OR_STR:
            if option in config:
                try:
                    value = config.getint(option)
                except ValueError:
                    value = config.get(option)
                setattr(cosmo, option, value)
        # biases: a file (read_* = 1), a constant (input_* present), or sampled values
        for kind, pops in (("galaxy", (1, 2, 3, 4)), ("magnification", (1, 2)), ("evolution", (1, 2))):
            for i in pops:
                name = f"{kind}_bias{i}"
                setter = getattr(cosmo, f"set_{name}")
                if f"read_{name}" in config and config.getint(f"read_{name}") == 1:
                    setter(*np.transpose(np.loadtxt(config.get(f"input_{name}").strip('" '))))
                elif f"input_{name}" in config:
                    setter(np.linspace(0, 10, 100), [config.getfloat(name)] * 100)
                elif f"{name}_redshifts" in config and f"{name}_values" in config:
                    setter(config.getfloat_array(f"{name}_redshifts"), config.getfloat_array(f"{name}_values"))
        return cosmo

    def _mbias_grid(self):
        """The four magnification / evolution biases on their common redshift samples."""
        x = self._mbias["s1"].x
        if not all(np.array_equal(x, self._mbias[k].x) for k in ("s2", "fevo1", "fevo2")):
            raise ValueError("the magnification and evolution biases must share their redshift sampling")
        return x, [self._mbias[k].y for k in ("s1", "s2", "fevo1", "fevo2")]

    # ---- background -------------------------------------------------------------
    def _background(self):
        if self._bg is None:
            mz, (s1, s2, e1, e2) = self._mbias_grid()
            self._bg = _background.compute(omega_cdm=self._p["omega_cdm"], omega_baryon=self._p["omega_baryon"],
                                           omega_gamma=self._p["omega_gamma"], w0=self._p["w0"], wa=self._p["wa"],
                                           h=self._p["h"], bins=self._p["background_sampling"], bias_z=mz,
                                           s1=s1, s2=s2, fevo1=e1, fevo2=e2)
        return self._bg

    def _bg_eval(self, key, z):
        _check_parameter("z", z, (int, float), 0, 15)
        bg = self._background()
        return float(_Spline(bg["z"], bg[key])(float(z))[0])

    def _bg_column(self, key, zs):
        """A background table at the redshifts zs in [0, 15] (one Akima pass)."""
        bg = self._background()
        return _Spline(bg["z"], bg[key])(np.asarray(zs, dtype=np.float64))

    def comoving_distance(self, z):
        return self._bg_eval("chi", z)

    def growth_factor(self, z):
        return self._bg_eval("D1", z)

    def growth_rate(self, z):
        return self._bg_eval("f", z)

    def hubble_rate(self, z):
        return self._bg_eval("Hz", z)

    def hubble_rate_conformal(self, z):
        return self._bg_eval("conformal_Hz", z)

    def hubble_rate_conformal_prime(self, z):
        return self._bg_eval("conformal_Hz_prime", z)

    def hubble_rate_conformal_derivative(self, z):
        """d(aH)/d tau at z, in 1/Mpc^2 (coffe.pyx:3090-3110)."""
        return self._bg_eval("conformal_Hz_prime", z)

    def scale_factor(self, z):
        return self._bg_eval("a", z)

    def _bg_tables(self):
        bg = self._background()
        return dict(z=bg["z"], chi=bg["chi"], D1=bg["D1"], f=bg["f"], cH=bg["conformal_Hz"], G1=bg["G1"],
                    G2=bg["G2"], a=bg["a"])

    # ---- integrals ----------------------------------------------------------------
    def _need_pk(self):
        if self._pk is None:
            raise ValueError("No power spectrum set: call set_power_spectrum_linear first")

    def _integral_tables(self):
        """The 8 default tables (integrals.c:580-660), each the reference's N + 43 nodes."""
        self._need_pk()
        if self._tables is None:
            from .engine import integral_table
            k, pk = self._pk
            N = self._p["integrals_fftlog_sampling"]
            kmin, kmax = self._kwin()
            self._tables = {t: integral_table(t[0], t[1], k, pk, N, kmin, kmax) for t in TABLES}
        return self._tables

    def _divergent_table(self):
        """I_0^4(r) r^4 at the reference's separations (integrals.c:772-834), the table the
        reference builds when d2 or g1-g5 is on; here built on first use."""
        self._need_pk()
        if self._divergent is None:
            from .engine import divergent_tables
            k, pk = self._pk
            kmin, kmax = self._kwin()
            sep, i40, r0, _, status, _ = divergent_tables(k, pk, kmin=kmin, kmax=kmax, return_status=True)
            if np.any(status[:len(sep)] == 1):
                raise RuntimeError("I_0^4: an integral reached the interval cap")
            self._divergent = (sep, i40)
            # the r = 0 renormalisation against separation, first node (0, 0) (integrals.c:829-835)
            x0, y0 = np.array(sep, dtype=np.float64), np.array(r0, dtype=np.float64)
            x0[0], y0[0] = 0.0, 0.0
            self._divergent_r0 = (x0, y0)
        return self._divergent

    def integral(self, *, r, n, l):  # noqa: E741
        """I_l^n(r) (times r^(n - l) when n > l), Akima of the table (coffe.pyx:1657-1713).
        (n, l) = (4, 0) is the divergent table, which the reference holds only when d2
        or g1-g5 is on (coffe.pyx:1108-1109, integrals.c:772-834)."""
        _check_parameter("r", r, (int, float), 0, 25000)
        _check_parameter("n", n, int, -2, 4)
        _check_parameter("l", l, int, 0, 4)
        if (n, l) == (4, 0):
            x, y = self._divergent_table()
            return float(_Spline(x, y)(float(r))[0])
        if (n, l) not in TABLES:
            raise ValueError(f"Cannot find integral with n = {n}, l = {l}")
        x, y = self._integral_tables()[(n, l)]
        return float(_Spline(x, y)(float(r))[0])

    # ---- multipoles ---------------------------------------------------------------
    def _check_coords_multipoles(self):
        if not (len(self._p["l"]) and len(self._p["sep"]) and len(self._p["z_mean"])):
            raise ValueError("At least one of the following parameters is empty: l, sep, z_mean")

    def _check_contributions(self):
        if not any(self._p[k] for k in ("has_density", "has_rsd", "has_d1", "has_d2", "has_g1", "has_g2", "has_g3",
                                        "has_g4", "has_g5", "has_lensing")):
            raise ValueError("No contributions specified, you need to specify at least one of: "
                             "'has_density', 'has_rsd', 'has_lensing, 'has_d1', 'has_d2', 'has_g1', 'has_g2', "
                             "'has_g3', 'has_g4', 'has_g5'")

    def _no_lensing(self, what):
        if self._p["has_lensing"] or self._p["has_g4"] or self._p["has_g5"]:
            raise NotImplementedError(f"the integrated terms (lensing, g4, g5) in {what} are not built")

    def _dg_sources(self):
        return {k for k in _DG if self._p["has_" + k]}

    def _terms(self, dg):
        """COFFE_TERM_* mask of the contribution flags; has_only_cross_correlations drops
        the auto terms (functions.c:129-669)."""
        from .engine import TERM_D1_D1, TERM_DEN_DEN, TERM_DG_MUTUAL, TERM_RSD_RSD, terms_of
        t = terms_of(self._p["has_density"], self._p["has_rsd"], self._p["has_d1"], dg is not None)
        if self._p["has_only_cross_correlations"]:
            t &= ~(TERM_DEN_DEN | TERM_RSD_RSD | TERM_D1_D1)
            if dg is not None:
                t |= TERM_DG_MUTUAL   # the d2-d2, g1-g1, ... terms left out, their mutual ones kept
        return t

    def _ren_grid(self, chi_max):
        """The 201-node renormalisation grid of I_0^4 on [0, chi_max] (integrals.c:848-907:
        chi_j = j / 200 chi_max), computed on the GPU for the current spectrum; cached."""
        from .engine import divergent_tables
        self._need_pk()
        if self._ren is None or self._ren[0] != chi_max:
            k, pk = self._pk
            chi = np.array([j / 200 * chi_max for j in range(201)])
            kmin, kmax = self._kwin()
            _, _, _, ren = divergent_tables(k, pk, sep=np.zeros(0), chi=chi, kmin=kmin, kmax=kmax)
            self._ren = (chi_max, chi, ren)
        return self._ren[1], self._ren[2]

    def _weights_setup(self, chi_max):
        """The tables the weights use and the d2 / g inputs (None without those terms)."""
        srcs = self._dg_sources()
        if not srcs:
            return list(TABLES), None
        mz, (s1, s2, e1, e2) = self._mbias_grid()
        chi, ren = self._ren_grid(chi_max)
        dg = dict(sources=srcs, H0h=COFFE_H0 * self._p["h"], Omega_m=self.omega_m, mb_z=mz, s1=s1, s2=s2, fevo1=e1,
                  fevo2=e2, ren_chi=chi, ren=ren)
        return list(TABLES) + [(4, 0)], dg

    def _table_values(self, tables, seps):
        """I_t(sep) for the tables (the divergent (4, 0) one from its own table)."""
        from .engine import akima as _ak
        tabs = self._integral_tables()
        rows = []
        for t in tables:
            x, y = self._divergent_table() if t == (4, 0) else tabs[t]
            rows.append(_ak(x, y, seps))
        return np.array(rows)

    def _multipole_values(self, zs, seps, ls, rule=None):
        """[z][l][r] multipoles: weights on the GPU (mu quadrature), tables at the separations
        (with the d2 / g terms also the renormalisation grid's part, on the GPU).  rule:
        an explicit (mu nodes, weights) rule in place of the Gauss-Legendre one."""
        import torch
        from .engine import mu_weights
        bz = self._bias[1].x
        if not np.array_equal(bz, self._bias[2].x):
            raise ValueError("galaxy_bias1 and galaxy_bias2 must share their redshift sampling")
        # the grid of the d2 / g terms spans chi(max(z_mean + deltaz)) (integrals.c:862-871)
        zmax = max(float(z) + float(dz) for z, dz in zip(self._p["z_mean"], self._p["deltaz"]))
        tables, dg = self._weights_setup(self.comoving_distance(zmax))
        terms = self._terms(dg)
        if not terms:
            return np.zeros((len(zs), len(ls), len(seps)))
        tv = self._table_values(tables, seps)
        out = np.empty((len(zs), len(ls), len(seps)))
        # the weights kernel takes at most 4 multipoles per launch
        for l0 in range(0, len(ls), 4):
            lc = [int(v) for v in ls[l0:l0 + 4]]
            ren = torch.empty((len(zs), len(lc), len(seps)), dtype=torch.float64, device="cuda") if dg else None
            A = mu_weights(self._engine_handle(), self._bg_tables(), bz, self._bias[1].y, self._bias[2].y, zs, lc,
                           tables, seps, flatsky=self._p["has_flatsky_local"],
                           nmu=self._p["multipole_quadrature_nodes"], terms=terms, dg=dg, ren_out=ren,
                           rule=rule).cpu().numpy()
            out[:, l0:l0 + len(lc)] = np.einsum("zltr,tr->zlr", A, tv)
            if ren is not None:
                out[:, l0:l0 + len(lc)] += ren.cpu().numpy()
        return out

    def multipoles_cosmo_jvp(self, direction):
        """d/d theta of compute_multipoles_bulk's values along `direction`, a dict over
        the background parameters omega_cdm, omega_baryon, omega_gamma, w0, wa, h (no
        reference counterpart): the exact tangents of the background tables
        (coffe_background_jvp) through the engine's forward-mode weights
        (coffe_mu_weights_jvp), full or flat sky, density / RSD / Doppler-1, at fixed P(k)
        and separations.  The six parameters are independent coordinates here: a step in
        omega_baryon holds omega_cdm (the omega_baryon setter holds omega_m instead).
        Returns an array in the records' order (z, l, r)."""
        from . import background as _bgm
        from .engine import mu_weights_jvp
        self._check_coords_multipoles()
        self._check_contributions()
        self._no_lensing("the cosmological tangent")
        if self._dg_sources():
            raise NotImplementedError("the cosmological tangent of the d2 / g1-g3 terms is not built")
        bz = self._bias[1].x
        if not np.array_equal(bz, self._bias[2].x):
            raise ValueError("galaxy_bias1 and galaxy_bias2 must share their redshift sampling")
        mz, (s1, s2, e1, e2) = self._mbias_grid()
        bg, tg = _bgm.jvp(direction, omega_cdm=self._p["omega_cdm"], omega_baryon=self._p["omega_baryon"],
                          omega_gamma=self._p["omega_gamma"], w0=self._p["w0"], wa=self._p["wa"], h=self._p["h"],
                          bias_z=mz, s1=s1, s2=s2, fevo1=e1, fevo2=e2)
        tabs = dict(z=bg["z"], chi=bg["chi"], D1=bg["D1"], f=bg["f"], cH=bg["conformal_Hz"], G1=bg["G1"], G2=bg["G2"])
        tans = dict(chi=tg["chi"], D1=tg["D1"], f=tg["f"], cH=tg["conformal_Hz"], G1=tg["G1"], G2=tg["G2"])
        zs, ls, seps = self._p["z_mean"], [int(v) for v in self._p["l"]], self._p["sep"]
        tables = list(TABLES)
        terms = self._terms(None)
        out = np.zeros((len(zs), len(ls), len(seps)))
        if terms:
            tv = self._table_values(tables, seps)
            for l0 in range(0, len(ls), 4):
                lc = ls[l0:l0 + 4]
                _, At = mu_weights_jvp(self._engine_handle(), tabs, tans, bz, self._bias[1].y, self._bias[2].y, zs, lc,
                                       tables, seps, flatsky=self._p["has_flatsky_local"],
                                       nmu=self._p["multipole_quadrature_nodes"], terms=terms)
                out[:, l0:l0 + len(lc)] = np.einsum("zltr,tr->zlr", At.cpu().numpy(), tv)
        return out.reshape(-1)

    def _lensing_tables(self, ls, want_ll, want_dl):
        """The (1, l) tables of the lensing-lensing term (integrals.c:1082-1130) and the
        half-integer tables ((l - 2k + 3)/2, (l - 2k + 1)/2) of the density-lensing term
        (integrals.c:1246-1300), for the even multipoles; cached."""
        from .engine import FL_MAXK, half_integer_table, integral_table
        key = (tuple(ls), bool(want_ll), bool(want_dl))
        if self._lens is None or self._lens[0] != key:
            self._need_pk()
            k, pk = self._pk
            N = self._p["integrals_fftlog_sampling"]
            kmin, kmax = self._kwin()
            even = sorted({int(v) for v in ls if int(v) % 2 == 0})
            if want_ll and even and even[-1] > 4:
                raise ValueError("the flat-sky lensing-lensing multipoles go up to l = 4 here")
            if want_dl and even and even[-1] > 2 * (FL_MAXK - 1):
                raise ValueError(f"the flat-sky density-lensing multipoles go up to l = {2 * (FL_MAXK - 1)} here")
            ll = {l: integral_table(1, l, k, pk, N, kmin, kmax) for l in even} if want_ll else {}
            dl = ({(l, j): half_integer_table(l - 2 * j + 3, l - 2 * j + 1, k, pk, N, kmin, kmax)
                   for l in even for j in range(l // 2 + 1)} if want_dl else {})
            self._lens = (key, ll, dl)
        return self._lens[1], self._lens[2]

    def _lensing_plan(self):
        """How the integrated terms are computed (signal.c:556-771), or None without
        lensing, g4 and g5: a dict with
          flat_ll  the flat-sky lensing-lensing term (has_flatsky_nonlocal; an auto term,
                   so not with has_only_cross_correlations),
          flat_dl  the flat-sky density-lensing term (has_flatsky_local_nonlocal; RSD- and
                   Doppler-lensing vanish in flat sky),
          flags    the contributions of the full-sky integrated terms (coffe_integrated_
                   corrfunc), with flat_ll / flat_dl marking what the flat-sky forms
                   replace, or None when none remains."""
        p = self._p
        lens, g4, g5 = p["has_lensing"], p["has_g4"], p["has_g5"]
        if not (lens or g4 or g5):
            return None
        oc = p["has_only_cross_correlations"]
        local = [k for k in ("density", "rsd", "d1", "d2", "g1", "g2", "g3") if p["has_" + k]]
        names = {"density": "den"}
        flags = {names.get(k, k) for k in local}
        if lens:
            flags.add("len")
        if g4:
            flags.add("g4")
        if g5:
            flags.add("g5")
        if oc:
            flags.add("only_cross")
        if p["has_flatsky_nonlocal"]:
            flags.add("flat_ll")
        if p["has_flatsky_local_nonlocal"]:
            flags.add("flat_dl")
        dl_local = bool(set(local) & {"density", "rsd", "d1"}) and not p["has_flatsky_local_nonlocal"]
        double = (lens and not oc and not p["has_flatsky_nonlocal"]) or ((g4 or g5) and not oc) or \
            (lens and (g4 or g5)) or (g4 and g5)
        single = (lens and (dl_local or bool(set(local) & {"d2", "g1", "g2", "g3"}))) or ((g4 or g5) and bool(local))
        return dict(flat_ll=bool(lens and not oc and p["has_flatsky_nonlocal"]),
                    flat_dl=bool(lens and p["has_density"] and p["has_flatsky_local_nonlocal"]),
                    flags=flags if (double or single) else None)

    def _lensing_values(self, zs, seps, ls):
        """[z][l][r] flat-sky lensing multipoles on the GPU (coffe_flat_lensing_multipoles),
        or None when neither flat-sky term is on."""
        from .engine import flat_lensing
        plan = self._lensing_plan()
        if plan is None or not (plan["flat_ll"] or plan["flat_dl"]):
            return None
        want_ll, want_dl = plan["flat_ll"], plan["flat_dl"]
        ll, dl = self._lensing_tables(ls, want_ll, want_dl)
        zpar = [[float(z), self.comoving_distance(float(z)), self.growth_factor(float(z)),
                 self._eval_mbias("s1", float(z)), self._eval_mbias("s2", float(z)),
                 self._eval_bias(1, float(z)), self._eval_bias(2, float(z))] for z in zs]
        return flat_lensing(zpar, [int(v) for v in ls], seps, self._background(), ll, dl, self.omega_m,
                            COFFE_H0 * self._p["h"], want_ll, want_dl)

    # theta-Gauss-Legendre nodes (mu = cos(theta)) of the integrated terms' multipoles
    INTEGRATED_MU_NODES = 64

    def _integrated_corrfunc(self, zs, seps, mus, chi_max=None):
        """[z][mu][r] single- plus double-integrated correlation function, full sky
        (coffe_integrated_corrfunc), or None when the plan has no full-sky terms.  chi_max:
        the renormalisation grid's extent (default chi(max(z_mean + deltaz)))."""
        from .engine import integrated_corrfunc
        plan = self._lensing_plan()
        if plan is None or plan["flags"] is None:
            return None
        flags = plan["flags"]
        bz = self._bias[1].x
        if not np.array_equal(bz, self._bias[2].x):
            raise ValueError("galaxy_bias1 and galaxy_bias2 must share their redshift sampling")
        mz, (s1, s2, e1, e2) = self._mbias_grid()
        tables = dict(self._integral_tables())
        ren = ren0 = None
        if flags & {"g4", "g5"}:
            tables[(4, 0)] = self._divergent_table()
            ren0 = self._divergent_r0
            if chi_max is None:
                # the grid spans chi(max(z_mean + deltaz)) (integrals.c:862-871)
                zmax = max(float(z) + float(dz) for z, dz in zip(self._p["z_mean"], self._p["deltaz"]))
                chi_max = self.comoving_distance(zmax)
            ren = self._ren_grid(chi_max)
        single, double = integrated_corrfunc(self._bg_tables(), bz, self._bias[1].y, self._bias[2].y, mz, s1, s2, e1,
                                             e2, self.omega_m, COFFE_H0 * self._p["h"], tables, flags, zs, mus, seps,
                                             ren=ren, ren0=ren0)
        return single + double

    def _integrated_xi(self, zs, seps, mus, chi_max=None):
        """[z][mu][r] every integrated term of the correlation function (the flat-sky
        lensing forms, coffe_flat_lensing_corrfunc, and the full-sky integrals), or None
        without lensing, g4 and g5."""
        from .engine import flat_lensing_corrfunc, half_integer_table
        plan = self._lensing_plan()
        if plan is None:
            return None
        out = np.zeros((len(zs), len(mus), len(seps)))
        if plan["flat_ll"] or plan["flat_dl"]:
            # the (1/2, -1/2) table (integrals.c:917-962), cached with the other lensing tables
            if self._lens_cf is None:
                k, pk = self._pk
                kmin, kmax = self._kwin()
                self._lens_cf = half_integer_table(1, -1, k, pk, self._p["integrals_fftlog_sampling"], kmin, kmax)
            zpar = [[float(z), self.comoving_distance(float(z)), self.growth_factor(float(z)),
                     self._eval_mbias("s1", float(z)), self._eval_mbias("s2", float(z)),
                     self._eval_bias(1, float(z)), self._eval_bias(2, float(z))] for z in zs]
            out += flat_lensing_corrfunc(zpar, [float(m) for m in mus], seps, self._background(), self._lens_cf,
                                         self.omega_m, COFFE_H0 * self._p["h"], plan["flat_ll"], plan["flat_dl"])
        xi = self._integrated_corrfunc(zs, seps, [float(m) for m in mus], chi_max)
        if xi is not None:
            out += xi
        return out

    def _integrated_multipoles(self, zs, seps, ls):
        """[z][l][r] multipoles of the full-sky integrated terms: (2 l + 1) / 2 int dmu
        P_l(mu) xi(mu) (signal.c:143-230, 664-672, 743-752), mu = cos(theta) with theta on
        an INTEGRATED_MU_NODES-point Gauss-Legendre rule; None without such terms."""
        t, w = np.polynomial.legendre.leggauss(self.INTEGRATED_MU_NODES)
        theta = np.pi / 2 * (t + 1)
        mus, wm = np.cos(theta), np.pi / 2 * w * np.sin(theta)
        xi = self._integrated_corrfunc(zs, seps, mus)
        if xi is None:
            return None
        out = np.empty((len(zs), len(ls), len(seps)))
        for il, l in enumerate(ls):
            pl = np.polynomial.legendre.legval(mus, [0] * int(l) + [1])
            out[:, il, :] = (2 * int(l) + 1) / 2 * np.einsum("m,zmr->zr", wm * pl, xi)
        return out

    def _all_multipoles(self, zs, seps, ls):
        """The local terms and the integrated ones (flat-sky forms and full-sky integrals)."""
        self._lensing_plan()
        vals = self._multipole_values(zs, seps, ls) if self._has_local() else np.zeros((len(zs), len(ls), len(seps)))
        for extra in (self._lensing_values(zs, seps, ls), self._integrated_multipoles(zs, seps, ls)):
            if extra is not None:
                vals = vals + extra
        return vals

    def _has_local(self):
        return any(self._p[k] for k in ("has_density", "has_rsd", "has_d1", "has_d2", "has_g1", "has_g2", "has_g3"))

    def compute_multipole(self, *, z, r, l, recompute=False):  # noqa: E741
        """The multipole at (z_mean, r, l) (coffe.pyx:3232-3300)."""
        self._check_coords_multipoles()
        self._check_contributions()
        _check_parameter("z", z, (int, float), 0, 15)
        _check_parameter("r", r, (int, float), 0, 1500)
        _check_parameter("l", l, int, 0, 10)
        if recompute:
            self._free_all()
        return float(self._all_multipoles([float(z)], [float(r)], [int(l)])[0, 0, 0])

    def compute_multipoles_bulk(self, recompute=False):
        """Multipoles at every (z_mean, l, sep), z-major then l then r (multipoles.c:76-86)."""
        self._check_coords_multipoles()
        self._check_contributions()
        if recompute:
            self._free_all()
        if self._multipoles is None:
            zs, ls, seps = self._p["z_mean"], self._p["l"], self._p["sep"]
            vals = self._all_multipoles(zs, seps, ls)
            self._multipoles = np.array([Multipoles(z=float(z), r=float(r), l=int(li), value=float(vals[iz, il, ir]))
                                         for iz, z in enumerate(zs) for il, li in enumerate(ls)
                                         for ir, r in enumerate(seps)])
        return self._multipoles

    # ---- correlation function -----------------------------------------------------
    def _check_coords_corrfunc(self):
        if not (len(self._p["mu"]) and len(self._p["sep"]) and len(self._p["z_mean"])):
            raise ValueError("At least one of the following parameters is empty: mu, sep, z_mean")

    def _corrfunc_values(self, zs, seps, mus):
        """[z][mu][r] correlation function: the local terms (functions_nonintegrated,
        functions.c:89-1484, at each mu: the weights of a one-node rule, the node mu with
        weight 2 and l = 0), the flat-sky lensing terms (coffe_flat_lensing_corrfunc) and
        the full-sky integrated ones (coffe_integrated_corrfunc)."""
        self._lensing_plan()
        out = np.zeros((len(zs), len(mus), len(seps)))
        if self._has_local():
            for j, mu in enumerate(mus):
                out[:, j, :] = self._multipole_values(zs, seps, [0], rule=([float(mu)], [2.0]))[:, 0, :]
        xi = self._integrated_xi(zs, seps, [float(m) for m in mus])
        if xi is not None:
            out += xi
        return out

    def compute_corrfunc(self, *, z, r, mu, recompute=False):
        """The correlation function at (z_mean, r, mu) (coffe.pyx:3159-3230)."""
        self._check_coords_corrfunc()
        self._check_contributions()
        self._lensing_plan()
        _check_parameter("z", z, (int, float), 0, 15)
        _check_parameter("r", r, (int, float), 0, 1500)
        _check_parameter("mu", mu, (int, float), -1, 1)
        if recompute:
            self._free_all()
        return float(self._corrfunc_values([float(z)], [float(r)], [float(mu)])[0, 0, 0])

    def compute_corrfunc_bulk(self, recompute=False):
        """The correlation function at every (z_mean, mu, sep), z-major then mu then r
        (corrfunc.c:83-90)."""
        self._check_coords_corrfunc()
        self._check_contributions()
        self._lensing_plan()
        if recompute:
            self._free_all()
        if self._corrfunc is None:
            zs, mus, seps = self._p["z_mean"], self._p["mu"], self._p["sep"]
            vals = self._corrfunc_values(zs, seps, mus)
            self._corrfunc = np.array([Corrfunc(z=float(z), r=float(r), mu=float(mu), value=float(vals[iz, im, ir]))
                                       for iz, z in enumerate(zs) for im, mu in enumerate(mus)
                                       for ir, r in enumerate(seps)])
        return self._corrfunc

    def maximum_separation(self, z_mean, deltaz):
        """The largest separation of the bin [z_mean - deltaz, z_mean + deltaz]: 2 (chi(z_mean
        + deltaz) - chi(z_mean)) (coffe.pyx:2764-2792)."""
        _check_parameter("z_mean", z_mean, (int, float), 0, 15)
        _check_parameter("deltaz", deltaz, (int, float), 0, z_mean)
        return 2 * (self.comoving_distance(z_mean + deltaz) - self.comoving_distance(z_mean))

    # ---- redshift-averaged multipoles ----------------------------------------------
    def _check_coords_average_multipoles(self):
        zmin, zmax = self._p["z_min"], self._p["z_max"]
        if not (len(self._p["l"]) and len(self._p["sep"]) and len(zmin)):
            raise ValueError("At least one of the following parameters is empty: l, sep, z_min")
        if len(zmin) != len(zmax):
            raise ValueError("z_min and z_max must have the same length")
        if np.any(np.asarray(zmin) >= np.asarray(zmax)):
            raise ValueError("every z_min must be below its z_max")

    def _average_values(self, zmin, zmax, seps, ls, nodes=32):
        """[l][r] multipoles averaged over the redshift bin (signal.c:249-311, 543-550): for
        each separation r, z runs over [z1, z2] with z1 = z(chi(zmin) + r/2) and
        z2 = z(chi(zmax) - r/2), and the average is int_0^1 dv M_l(z(v), r) h H0 / H(z(v)),
        z(v) = z1 + (z2 - z1) v (the reference's integrand, no further normalisation); the
        v integral is a Gauss-Legendre rule of `nodes` points, the M_l(z, r) the engine's
        weights (GPU mu quadrature) at exactly the (z(v), r) node pairs (paired mode)."""
        import torch
        from .engine import mu_weights
        bg = self._background()
        z_of_chi = _Spline(bg["chi"], bg["z"])
        chi_of_z = _Spline(bg["z"], bg["chi"])
        hz = _Spline(bg["z"], bg["Hz"])
        x, w = np.polynomial.legendre.leggauss(nodes)
        v, w = 0.5 * (x + 1), 0.5 * w
        seps = np.asarray(seps, dtype=np.float64)
        z1 = z_of_chi(chi_of_z(float(zmin))[0] + seps / 2)
        z2 = z_of_chi(chi_of_z(float(zmax))[0] - seps / 2)
        if np.any(z2 <= z1):
            raise ValueError("a separation exceeds the comoving depth of the redshift bin")
        zn = z1[:, None] + (z2 - z1)[:, None] * v[None, :]                  # [r][v]
        hH0 = self._p["h"] * COFFE_H0
        wz = w[None, :] * hH0 / hz(zn.ravel()).reshape(zn.shape)           # [r][v]
        bz = self._bias[1].x
        if not np.array_equal(bz, self._bias[2].x):
            raise ValueError("galaxy_bias1 and galaxy_bias2 must share their redshift sampling")
        # the d2 / g grid spans chi(max(z_max)) for averaged multipoles (integrals.c:873-879)
        chi_max = self.comoving_distance(float(max(self._p["z_max"])))
        out = np.zeros((len(ls), len(seps)))
        # the integrated terms (signal.c:312-470, 674-682, 759-767): their correlation
        # function at the (z(v), mu) nodes of every separation, projected on P_l and
        # averaged with the same weights
        if self._lensing_plan() is not None:
            t_, w_ = np.polynomial.legendre.leggauss(self.INTEGRATED_MU_NODES)
            theta = np.pi / 2 * (t_ + 1)
            mus, wm = np.cos(theta), np.pi / 2 * w_ * np.sin(theta)
            pls = np.array([np.polynomial.legendre.legval(mus, [0] * int(l) + [1]) * (2 * int(l) + 1) / 2
                            for l in ls])
            for ir, r in enumerate(seps):
                xi = self._integrated_xi(zn[ir], [float(r)], mus, chi_max)[:, :, 0]    # [v][mu]
                out[:, ir] += np.einsum("v,vm,lm->l", wz[ir], xi * wm[None, :], pls)
        if not self._has_local():
            return out
        tables, dg = self._weights_setup(chi_max)
        npair = len(seps) * nodes
        terms = self._terms(dg)
        if not terms:
            return out
        tv = self._table_values(tables, seps)
        for l0 in range(0, len(ls), 4):   # at most 4 multipoles per weights launch
            lc = [int(v) for v in ls[l0:l0 + 4]]
            ren = torch.empty((npair, len(lc)), dtype=torch.float64, device="cuda") if dg else None
            A = mu_weights(self._engine_handle(), self._bg_tables(), bz, self._bias[1].y, self._bias[2].y,
                           zn.ravel(), lc, tables, np.repeat(seps, nodes), flatsky=self._p["has_flatsky_local"],
                           nmu=self._p["multipole_quadrature_nodes"], terms=terms, dg=dg, ren_out=ren,
                           paired=True).cpu().numpy()
            A = A.reshape(len(seps), nodes, len(lc), len(tables))                 # [r][v][l][t]
            out[l0:l0 + len(lc)] += np.einsum("rv,rvlt,tr->lr", wz, A, tv)
            if ren is not None:
                out[l0:l0 + len(lc)] += np.einsum("rv,rvl->lr", wz, ren.cpu().numpy().reshape(len(seps), nodes,
                                                                                              len(lc)))
        return out

    def compute_average_multipoles_bulk(self, recompute=False):
        """Redshift-averaged multipoles at every (z_min/z_max bin, l, sep), bin-major then
        l then r (average_multipoles.c:90-146)."""
        self._check_coords_average_multipoles()
        self._check_contributions()
        if recompute:
            self._free_all()
        if self._average_multipoles is None:
            ls, seps = self._p["l"], self._p["sep"]
            out = []
            for zmin, zmax in zip(self._p["z_min"], self._p["z_max"]):
                vals = self._average_values(zmin, zmax, seps, ls)
                out += [AverageMultipoles(z_min=float(zmin), z_max=float(zmax), r=float(r), l=int(li),
                                          value=float(vals[il, ir]))
                        for il, li in enumerate(ls) for ir, r in enumerate(seps)]
            self._average_multipoles = np.array(out)
        return self._average_multipoles

    # ---- covariance ---------------------------------------------------------------
    def _covariance_values(self, zs, vols):
        """cov[bin][l1][r1][l2][r2]: the 2D FFTLog integrals of P and P^2 and the Gaussian
        assembly (covariance.c:477-1103) with the biases, f and D1 at each bin's redshift
        zs[bin] and the volumes vols[bin]."""
        import torch
        from .covariance import CovPlan
        if self._p["covariance_interpolation_method"] != 2:
            raise NotImplementedError("the covariance is interpolated bicubically here "
                                      "(covariance_interpolation_method 'bicubic')")
        k, pk = self._pk
        ls, seps = [int(v) for v in self._p["l"]], self._p["sep"]
        n1 = np.asarray(self._p["number_density1"], dtype=np.float64)
        n2 = np.asarray(self._p["number_density2"], dtype=np.float64)
        if np.any(n1 <= 0) or np.any(n2 <= 0):
            raise ValueError("number_density1 and number_density2 must be positive for the covariance")
        kmin, kmax = self._kwin()
        plan = CovPlan(self._engine_handle(), self._p["covariance_integration_sampling"], kmin, kmax, ls, seps)
        try:
            integ = plan.integrals(torch.tensor(k, device="cuda"), torch.tensor(pk, device="cuda"), [len(k)])
            # galaxy biases 1..4 only with a density contribution, f only with RSD
            # (covariance.c:842-860); the populations pick which of them each field reads
            bias = [[self._eval_bias(i, float(z)) if self._p["has_density"] else 0.0 for i in (1, 2, 3, 4)]
                    for z in zs]
            f = [self.growth_rate(float(z)) if self._p["has_rsd"] else 0.0 for z in zs]
            D1 = [self.growth_factor(float(z)) for z in zs]
            return plan.assemble(integ, np.array(bias), f, D1, n1, list(vols), self._p["pixelsize"],
                                 poisson=self._p["covariance_poisson"], mixed=self._p["covariance_mixed"],
                                 cosmic=self._p["covariance_cosmic"], pops=self._p["covariance_populations"],
                                 nbar2=n2, window=self._p["has_binned_covariance"]).cpu().numpy()[0]
        finally:
            plan.close()

    def compute_covariance_bulk(self, recompute=False):
        """Gaussian flat-sky covariance of the multipoles with the 2D FFTLog
        (covariance.c:477-1103); records ordered z, l1, r1, l2, r2.  The volume of bin k is
        the shell 4 pi fsky (chi(z + deltaz)^3 - chi(z - deltaz)^3) / 3 (covariance.c:
        793-806)."""
        self._check_coords_multipoles()
        self._check_contributions()
        self._need_pk()
        zs = self._p["z_mean"]
        per_bin = [self._p[k] for k in ("deltaz", "number_density1", "number_density2", "pixelsize", "fsky")]
        if any(len(v) != len(zs) for v in per_bin):
            raise ValueError("Mismatching lengths for covariance parameters (`z_mean`, `deltaz`, `number_density1`, "
                             "`number_density2`, `pixelsize`, `fsky`")
        if recompute:
            self._free_all()
        if self._covariance is None:
            ls, seps = [int(v) for v in self._p["l"]], self._p["sep"]
            vol = []
            for iz, z in enumerate(zs):
                dz, fsky = self._p["deltaz"][iz], self._p["fsky"][iz]
                chi_lo = self.comoving_distance(float(z - dz))
                chi_hi = self.comoving_distance(float(z + dz))
                vol.append(4 * np.pi * fsky * (chi_hi ** 3 - chi_lo ** 3) / 3.)
            cov = self._covariance_values(zs, vol)
            self._covariance = np.array([
                Covariance(r1=float(seps[m]), r2=float(seps[n]), l1=ls[i], l2=ls[j], z=float(z),
                           value=float(cov[iz, i, m, j, n]))
                for iz, z in enumerate(zs) for i in range(len(ls)) for m in range(len(seps))
                for j in range(len(ls)) for n in range(len(seps))])
        return self._covariance

    def _average_volume(self, zmin, zmax, fsky):
        """4 pi fsky / (int_zmin^zmax dz / (H chi^2 (1 + z)) (H0 h)^2), H the conformal
        Hubble rate (covariance.c:150-162, 808-837)."""
        from scipy.integrate import quad
        bg = self._background()
        chi, cH = _Spline(bg["z"], bg["chi"]), _Spline(bg["z"], bg["conformal_Hz"])
        integ = quad(lambda z: 1.0 / cH(z)[0] / chi(z)[0] ** 2 / (1 + z), zmin, zmax, epsabs=0, epsrel=1e-12,
                     limit=200)[0]
        hH0 = self._p["h"] * COFFE_H0
        return 4 * np.pi * fsky / integ / hH0 ** 2

    def compute_average_covariance_bulk(self, recompute=False):
        """Gaussian covariance of the redshift-averaged multipoles (covariance.c:477-1103,
        COVARIANCE_AVERAGE_MULTIPOLES; coffe.pyx:3482-3535): per bin [z_min, z_max] the
        flat-sky covariance at z = (z_min + z_max)/2 with the volume of _average_volume;
        records ordered bin, l1, r1, l2, r2."""
        self._check_coords_multipoles()
        self._check_contributions()
        self._need_pk()
        zmin, zmax = self._p["z_min"], self._p["z_max"]
        per_bin = [self._p[k] for k in ("deltaz", "z_max", "number_density1", "number_density2", "pixelsize",
                                        "fsky")]
        if any(len(v) != len(zmin) for v in per_bin):
            raise ValueError("Mismatching lengths for average covariance parameters (`z_min`, `z_max`, `deltaz`, "
                             "`number_density1`, `number_density2`, `pixelsize`, `fsky`")
        for a, b in zip(zmin, zmax):
            self.maximum_separation((a + b) / 2, (b - a) / 2)
        if recompute:
            self._free_all()
        if self._average_covariance is None:
            ls, seps = [int(v) for v in self._p["l"]], self._p["sep"]
            zs = [(a + b) / 2 for a, b in zip(zmin, zmax)]
            vol = [self._average_volume(float(a), float(b), float(fs))
                   for a, b, fs in zip(zmin, zmax, self._p["fsky"])]
            cov = self._covariance_values(zs, vol)
            self._average_covariance = np.array([
                AverageCovariance(r1=float(seps[m]), r2=float(seps[n]), l1=ls[i], l2=ls[j], z_min=float(a),
                                  z_max=float(b), value=float(cov[iz, i, m, j, n]))
                for iz, (a, b) in enumerate(zip(zmin, zmax)) for i in range(len(ls)) for m in range(len(seps))
                for j in range(len(ls)) for n in range(len(seps))])
        return self._average_covariance


def _kept_property(name, kind, lo, hi):
    def setter(self, value):
        _check_parameter(name, value, kind, lo, hi)
        self._p[name] = value
    return property(lambda s: s._p[name], setter)


for _name, (_kind, _lo, _hi) in _KEPT.items():
    setattr(Coffe, _name, _kept_property(_name, _kind, _lo, _hi))
