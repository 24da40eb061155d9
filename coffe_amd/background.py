"""Background tables (coffe_background_compute in the engine library) and their exact
derivatives with respect to the cosmological parameters (coffe_background_jvp)."""
import ctypes

import numpy as np

from ._lib import check, declare_background, load_library

NAMES = ["z", "a", "Hz", "conformal_Hz", "conformal_Hz_prime", "D1", "f", "chi", "G1", "G2"]


def compute(omega_cdm=0.25, omega_baryon=0.05, omega_gamma=9e-5, w0=-1.0, wa=0.0, h=0.67, bins=10000,
            bias_z=None, s1=None, s2=None, fevo1=None, fevo2=None):
    L = declare_background(load_library())
    if bias_z is None:
        bias_z = np.arange(11.0)
    bias_z = np.ascontiguousarray(bias_z, dtype=np.float64)
    zeros = np.zeros_like(bias_z)
    arrs = [np.ascontiguousarray(v if v is not None else zeros, dtype=np.float64) for v in (s1, s2, fevo1, fevo2)]
    outs = [np.empty(bins) for _ in NAMES]
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    check(L.coffe_background_compute(omega_cdm, omega_baryon, omega_gamma, w0, wa, h, bins, len(bias_z), dp(bias_z),
                                     *[dp(a) for a in arrs], *[dp(o) for o in outs]), "coffe_background_compute")
    return dict(zip(NAMES, outs))


PARAMS = ["omega_cdm", "omega_baryon", "omega_gamma", "w0", "wa", "h"]


def jvp(direction, omega_cdm=0.25, omega_baryon=0.05, omega_gamma=9e-5, w0=-1.0, wa=0.0, h=0.67, bins=10000,
        bias_z=None, s1=None, s2=None, fevo1=None, fevo2=None):
    """(tables, tangents): the tables of `compute` and their derivatives along
    `direction`, a dict {parameter name: component} over PARAMS (missing names are 0),
    e.g. {"omega_cdm": 1.0} for d/d Omega_cdm at fixed Omega_baryon."""
    L = declare_background(load_library())
    unknown = set(direction) - set(PARAMS)
    if unknown:
        raise ValueError(f"unknown parameters {sorted(unknown)}; expected some of {PARAMS}")
    par = np.array([omega_cdm, omega_baryon, omega_gamma, w0, wa, h], dtype=np.float64)
    dvec = np.array([float(direction.get(n, 0.0)) for n in PARAMS])
    if bias_z is None:
        bias_z = np.arange(11.0)
    bias_z = np.ascontiguousarray(bias_z, dtype=np.float64)
    zeros = np.zeros_like(bias_z)
    arrs = [np.ascontiguousarray(v if v is not None else zeros, dtype=np.float64) for v in (s1, s2, fevo1, fevo2)]
    out = np.empty((len(NAMES), bins))
    tout = np.empty((len(NAMES), bins))
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    check(L.coffe_background_jvp(dp(par), dp(dvec), bins, len(bias_z), dp(bias_z), *[dp(a) for a in arrs], dp(out),
                                 dp(tout)), "coffe_background_jvp")
    return dict(zip(NAMES, out)), dict(zip(NAMES, tout))
