"""The C-ABI library loads and exports every function declared in include/*.h
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", src, flags=re.M):
            if not m.group(0).lstrip().startswith("typedef"):
                names.add(m.group(1))
    return names


def test_library_exports_all_declared_symbols():
    from coffe_amd._lib import LIB_PATH
    assert os.path.exists(LIB_PATH), "engine library not built"
    lib = ctypes.CDLL(LIB_PATH)
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_error_strings_host_only():
    from coffe_amd._lib import load_library
    L = load_library()
    for code in range(9):
        assert L.coffe_engine_strerror(code)
    assert L.coffe_engine_last_error() == 0


class _GslSplineView:
    """ctypes mirrors of the reference's coffe_interpolation (common.h:228-232) and the
    head of GSL's gsl_spline (interp, x, y, size)."""

    @staticmethod
    def make(x, y):
        class Spline(ctypes.Structure):
            _fields_ = [("interp", ctypes.c_void_p), ("x", ctypes.POINTER(ctypes.c_double)),
                        ("y", ctypes.POINTER(ctypes.c_double)), ("size", ctypes.c_size_t)]

        class Interp(ctypes.Structure):
            _fields_ = [("spline", ctypes.POINTER(Spline)), ("accel", ctypes.c_void_p)]

        dp = ctypes.POINTER(ctypes.c_double)
        sp = Spline(None, x.ctypes.data_as(dp), y.ctypes.data_as(dp), len(x))
        return Interp(ctypes.pointer(sp), None), sp


def test_integrals_renormalizable_rejects_bad_states():
    """The drop-in of integrals.c:459-573 takes both states integer or both half-integer
    (the flat-sky lensing tables); mixed states return COFFE_EUNSUPPORTED and a
    half-integer state with an even raw value COFFE_EINVAL, before any device work and
    with nothing allocated."""
    import numpy as np
    from coffe_amd._lib import load_library
    L = load_library()
    k = np.logspace(-4, 1, 64)
    pk = k / (1 + k ** 2)
    interp, keep = _GslSplineView.make(k, pk)
    f = L.coffe_integrals_renormalizable
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
    EINVAL, EUNSUPPORTED = 1, 4
    for n, l, sn, sl, want in ((1, -1, 1, 0, EUNSUPPORTED), (1, 1, 0, 1, EUNSUPPORTED), (2, -1, 1, 1, EINVAL),
                               (1, 2, 1, 1, EINVAL), (1, -3, 1, 1, EINVAL)):
        ox, oy = ctypes.POINTER(ctypes.c_double)(), ctypes.POINTER(ctypes.c_double)()
        n_out = ctypes.c_size_t(0)
        rc = f(ctypes.byref(ox), ctypes.byref(oy), 1024, ctypes.byref(n_out), ctypes.byref(interp), n, l, sn, sl,
               float(k[0]), float(k[-1]))
        assert rc == want, (n, l, sn, sl, rc)
        assert not ox and not oy and n_out.value == 0
