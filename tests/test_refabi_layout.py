"""The reference-ABI layer (include/coffe_refabi.h) on the host: the structure layouts
against the reference's headers and the ctypes mirror, and the entry points' host-only
behaviour (look-ups, coefficients, flag gating and error codes; no device work)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import tests.refabi_host as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"
EUNSUPPORTED, EDOMAIN = 4, 5

STRUCTS = [("coffe_parameters_t", R.Parameters), ("coffe_background_t", R.Background),
           ("coffe_integral_t", R.Integral), ("coffe_multipoles_t", R.Multipoles),
           ("coffe_average_multipoles_t", R.AverageMultipoles), ("coffe_covariance_t", R.Covariance),
           ("coffe_average_covariance_t", R.AverageCovariance)]


def _offsets_from_c(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "coffe_refabi.h"', "int main(void) {"]
    for cname, cls in STRUCTS:
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            lines.append(f'printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["return 0; }"]
    src = tmp_path / "offsets.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "offsets"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {tuple(line.split()[:2]): int(line.split()[2]) for line in out.splitlines()}


def test_ctypes_mirror_matches_header(tmp_path):
    """Every field of the ctypes mirror sits at the header's offset (gcc's layout)."""
    c = _offsets_from_c(tmp_path)
    for cname, cls in STRUCTS:
        assert c[(cname, "sizeof")] == ctypes.sizeof(cls), cname
        for f in cls._fields_:
            assert c[(cname, f[0])] == getattr(cls, f[0]).offset, (cname, f[0])


def _fields(text, start, end):
    """declared names of a C struct body, in order (comments, preprocessor lines and
    array sizes dropped)"""
    body = text[text.index(start):]
    body = body[body.index("{") + 1: body.index(end)]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    body = re.sub(r"//[^\n]*", "", body)
    body = re.sub(r"^\s*#[^\n]*", "", body, flags=re.M)
    names = []
    for decl in body.split(";"):
        decl = re.sub(r"\[[^\]]*\]", "", decl.strip())
        if not decl:
            continue
        m = re.match(r"^(?:const\s+)?(?:enum\s+|struct\s+)?\w+\s*(.*)$", decl, re.S)
        names += [n.strip().lstrip("*").strip() for n in m.group(1).split(",")]
    return names


@pytest.mark.skipif(not os.path.isdir(REF), reason="the reference sources are not on this machine")
def test_header_field_order_matches_reference():
    """coffe_parameters_t, coffe_background_t and the integral / record structures of the
    header declare the reference's fields in the reference's order (the C build, with
    the libconfig pointer)."""
    mine = open(os.path.join(ROOT, "include", "coffe_refabi.h")).read()
    pairs = [("common.h", "typedef struct coffe_parameters_t", "} coffe_parameters_t;"),
             ("background.h", "typedef struct coffe_background_t", "} coffe_background_t;"),
             ("integrals.h", "typedef struct coffe_integral_t", "} coffe_integral_t;"),
             ("common.h", "typedef struct coffe_multipoles_coords_t", "} coffe_multipoles_coords_t;"),
             ("common.h", "typedef struct coffe_average_multipoles_coords_t", "} coffe_average_multipoles_coords_t;"),
             ("common.h", "typedef struct coffe_covariance_coords_t", "} coffe_covariance_coords_t;"),
             ("common.h", "typedef struct coffe_average_covariance_coords_t",
              "} coffe_average_covariance_coords_t;"),
             ("common.h", "typedef struct coffe_correlation_contributions", "} coffe_correlation_contributions;")]
    for hdr, start, end in pairs:
        ref = open(os.path.join(REF, hdr)).read()
        assert _fields(ref, start, end) == _fields(mine, start, end), start


@pytest.fixture
def lib():
    from coffe_amd._lib import load_library
    L = R.declare(load_library())
    yield L
    L.coffe_integrals_free(ctypes.byref(R.IntegralArray()))   # a good call clears the thread's error


def test_find_integral_and_coefficients(lib):
    arr = (R.Integral * 3)()
    for i, (n, l, s) in enumerate([(0, 0, 0), (1, 3, 0), (1, 3, 1)]):
        arr[i].n, arr[i].l, arr[i].state_n, arr[i].state_l = n, l, s, s
    ia = R.IntegralArray(arr, 3)
    base = ctypes.addressof(arr)
    for (n, l, s), idx in [((0, 0, 0), 0), ((1, 3, 0), 1), ((1, 3, 1), 2)]:
        got = lib.coffe_find_integral(ctypes.byref(ia), n, l, s, s)
        assert ctypes.addressof(got.contents) == base + idx * ctypes.sizeof(R.Integral)
    assert not lib.coffe_find_integral(ctypes.byref(ia), 2, 2, 0, 0)
    assert not lib.coffe_find_integral(ctypes.byref(R.IntegralArray(None, 0)), 0, 0, 0, 0)
    b1, b2, f = 1.3, 1.7, 0.8
    assert lib.covariance_coefficient(b1, b2, f, 0) == pytest.approx(b1 * b2 + (b1 + b2) * f / 3 + f * f / 5, 1e-15)
    assert lib.covariance_coefficient(b1, b2, f, 2) == pytest.approx(2 * (b1 + b2) * f / 3 + 4 * f * f / 7, 1e-15)
    assert lib.covariance_coefficient(b1, b2, f, 4) == pytest.approx(8 * f * f / 35, 1e-15)
    assert lib.covariance_coefficient(b1, b2, f, 6) == 0.0


def test_covariance_find(lib):
    recs = (R.Covariance * 4)()
    vals = [(1.0, 0, 2, 10.0, 20.0, 1.5), (1.0, 2, 2, 10.0, 20.0, 2.5), (1.0, 2, 2, 20.0, 10.0, 3.5),
            (0.5, 2, 2, 20.0, 10.0, 4.5)]
    for r, (z, l1, l2, s1, s2, v) in zip(recs, vals):
        r.coords.z_mean, r.coords.l1, r.coords.l2, r.coords.separation1, r.coords.separation2 = z, l1, l2, s1, s2
        r.value = v
    arr = R.CovarianceArray(recs, 4)
    assert lib.coffe_covariance_find(ctypes.byref(arr), 1.0, 2, 2, 20.0, 10.0).value == 3.5
    # matching within a relative 1e-3 (covariance.c:91-103)
    assert lib.coffe_covariance_find(ctypes.byref(arr), 0.5004, 2, 2, 20.01, 9.995).value == 4.5
    miss = lib.coffe_covariance_find(ctypes.byref(arr), 1.0, 4, 4, 10.0, 10.0)
    assert np.isnan(miss.value) and miss.coords.l1 == -1
    assert lib.coffe_engine_last_error() == EDOMAIN


def test_integrate_flag_gating(lib):
    """coffe_integrate returns 0 where the reference does (no contribution of the kind);
    the integrated terms need the background and tables (EINVAL without them, before
    any device work)."""
    host = R.Host(lib)
    try:
        par = R.Parameters()
        args = (None, None, 1.0, 0.9, 1.1, 50.0, 0.3, 0)
        assert lib.coffe_integrate(ctypes.byref(par), *args, R.NONINTEGRATED, R.MULTIPOLES) == 0.0
        assert lib.coffe_engine_last_error() == 0
        par.correlation_contrib.den = 1
        assert lib.coffe_integrate(ctypes.byref(par), *args, R.SINGLE_INTEGRATED, R.MULTIPOLES) == 0.0
        assert lib.coffe_integrate(ctypes.byref(par), *args, R.DOUBLE_INTEGRATED, R.MULTIPOLES) == 0.0
        assert lib.coffe_integrate(ctypes.byref(par), *args, R.NONINTEGRATED, R.COVARIANCE_MULTIPOLES) == 0.0
        par.correlation_contrib.len = 1
        assert np.isnan(lib.coffe_integrate(ctypes.byref(par), *args, R.SINGLE_INTEGRATED, R.MULTIPOLES))
        assert lib.coffe_engine_last_error() == 1
        assert np.isnan(lib.coffe_integrate(ctypes.byref(par), *args, R.DOUBLE_INTEGRATED, R.CORRFUNC))
        assert lib.coffe_engine_last_error() == 1
        assert np.isnan(lib.coffe_integrate(ctypes.byref(par), *args, R.DOUBLE_INTEGRATED, R.AVERAGE_MULTIPOLES))
        assert lib.coffe_engine_last_error() == 1
        # the derivative outputs change the local terms' form
        par.correlation_contrib.len = 0
        par.b_derivative = 1
        assert np.isnan(lib.coffe_integrate(ctypes.byref(par), *args, R.NONINTEGRATED, R.MULTIPOLES))
        assert lib.coffe_engine_last_error() == EUNSUPPORTED
        # g4 in a multipoles run, flat-sky lensing-lensing tables above l = 4, bilinear covariance
        # interpolation
        par.b_derivative = 0
        par.correlation_contrib.g4 = 1
        z = np.array([1.0])
        s = np.array([50.0])
        ls = np.array([0], dtype=np.int32)
        par.z_mean, par.z_mean_len = z.ctypes.data_as(R.DP), 1
        par.sep, par.sep_len = s.ctypes.data_as(R.DP), 1
        par.multipole_values, par.multipole_values_len = ls.ctypes.data_as(R.IP), 1
        mp = R.MultipolesArray()
        assert lib.coffe_multipoles_init(ctypes.byref(par), None, None, ctypes.byref(mp)) == 1
        assert mp.size == 0 and not mp.array
        par.zmin, par.zmin_len = z.ctypes.data_as(R.DP), 1
        par.zmax, par.zmax_len = z.ctypes.data_as(R.DP), 1
        ram = R.AverageMultipolesArray()
        assert lib.coffe_average_multipoles_init(ctypes.byref(par), None, None, ctypes.byref(ram)) == 1
        par.flatsky_nonlocal = 1
        par.output_type = R.MULTIPOLES
        l6 = np.array([0, 6], dtype=np.int32)
        par.multipole_values, par.multipole_values_len = l6.ctypes.data_as(R.IP), 2
        ia = R.IntegralArray()
        assert lib.coffe_integrals_init(ctypes.byref(par), None, ctypes.byref(ia)) == EUNSUPPORTED
        assert ia.size == 0 and not ia.array
        par.multipole_values, par.multipole_values_len = ls.ctypes.data_as(R.IP), 1
        par.output_type = R.COVARIANCE_MULTIPOLES
        par.covariance_integration_method = 1        # STANDARD: computed by FFTLog, accepted
        par.covariance_interpolation_method = 1      # bilinear: not offered
        cm, ca = R.CovarianceArray(), R.AverageCovarianceArray()
        assert lib.coffe_covariance_init(ctypes.byref(par), ctypes.byref(R.Background()), ctypes.byref(cm),
                                         ctypes.byref(ca)) == EUNSUPPORTED
        # another output type leaves the covariance arrays alone
        par.output_type = R.MULTIPOLES
        assert lib.coffe_covariance_init(ctypes.byref(par), None, ctypes.byref(cm), ctypes.byref(ca)) == 0
    finally:
        host.close()


def test_integrals_free_uses_host_routines(lib):
    """coffe_integrals_free hands every spline of every entry to the host's free routines
    (integrals.c:1426-1444) and empties the array; without host routines the layer
    reports EUNSUPPORTED instead of leaking silently."""
    host = R.Host(lib)
    try:
        arr = (R.Integral * 2)()
        for e in arr:
            e.result = host.spline([0, 1, 2, 3, 4], [0, 1, 4, 9, 16])
        # the array must come from malloc, as the library frees it: build it there
        libc = ctypes.CDLL(None)
        libc.malloc.restype = ctypes.c_void_p
        libc.malloc.argtypes = [ctypes.c_size_t]
        mem = libc.malloc(ctypes.sizeof(arr))
        ctypes.memmove(mem, ctypes.addressof(arr), ctypes.sizeof(arr))
        ia = R.IntegralArray(ctypes.cast(mem, ctypes.POINTER(R.Integral)), 2)
        before = len(host.objects)
        assert lib.coffe_integrals_free(ctypes.byref(ia)) == 0
        assert ia.size == 0 and not ia.array
        assert before - len(host.objects) == 2 and host.freed == 2
    finally:
        host.close()
    ia = R.IntegralArray(None, 1)
    assert lib.coffe_integrals_free(ctypes.byref(ia)) == 0   # nothing allocated: nothing to free
    keep = (R.Integral * 1)()
    ia = R.IntegralArray(keep, 1)
    assert lib.coffe_integrals_free(ctypes.byref(ia)) == EUNSUPPORTED   # no host routines registered
    assert ia.size == 1
