"""ctypes binding of libcoffe_engine.so (the C-ABI of include/coffe_engine.h)."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("COFFE_ENGINE_LIBRARY") or os.path.join(_HERE, "libcoffe_engine.so")

_lib = None

ERRORS = {
    0: "success", 1: "invalid argument", 2: "out of memory", 3: "HIP device error",
    4: "unsupported size", 5: "argument outside the domain", 6: "non-finite value", 7: "no HIP device",
    8: "adaptive quadrature did not converge",
}
COFFE_ENOCONV = 8


class EngineError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        msg = ERRORS.get(code, f"error {code}")
        if _lib is not None:
            try:
                msg = _lib.coffe_engine_strerror(code).decode()
            except Exception:  # pragma: no cover
                pass
        super().__init__(f"{what}: {msg} (code {code})" if what else f"{msg} (code {code})")


def check(code, what=""):
    if code != 0:
        raise EngineError(code, what)


def load_library():
    """Load the engine library; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    # one HIP runtime per process: with torch imported first the engine binds to the
    # libamdhip64.so.7 torch loaded (the same soname); loaded the other way round,
    # torch would run on the engine's runtime and find no device
    import torch  # noqa: F401
    if not os.path.exists(LIB_PATH):
        raise EngineError(7, f"{LIB_PATH} is missing; run __graft_entry__.build() or python coffe_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    vp, i, i64, d, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_size_t
    dp = ctypes.POINTER(ctypes.c_double)
    i64p = ctypes.POINTER(ctypes.c_int64)
    ip = ctypes.POINTER(ctypes.c_int)
    L.coffe_engine_strerror.argtypes = [i]
    L.coffe_engine_strerror.restype = ctypes.c_char_p
    L.coffe_engine_last_error.argtypes = []
    L.coffe_engine_create.argtypes = [i, ctypes.POINTER(vp)]
    L.coffe_engine_destroy.argtypes = [vp]
    L.coffe_engine_destroy.restype = None
    L.coffe_engine_inject_alloc_failure.argtypes = [i]
    L.coffe_engine_inject_alloc_failure.restype = None
    L.coffe_engine_default.argtypes = []
    L.coffe_engine_default.restype = vp
    L.coffe_engine_synchronize.argtypes = [vp]
    L.twofast_1bessel.argtypes = [dp, dp, sz, dp, dp, sz, d, d, d, d, d, d, ctypes.c_uint]
    L.twofast_1bessel.restype = None
    L.twofast_1bessel_status.argtypes = [dp, dp, sz, dp, dp, sz, d, d, d, d, d, d]
    L.twofast_1bessel_noalloc.argtypes = [vp, vp, sz, vp, vp, sz, d, d, d, d, d, d, vp]
    L.coffe_batch_create.argtypes = [vp, i, i64p, i64p, dp, dp, i, ip, ip, i, dp, i, i, ctypes.POINTER(vp)]
    L.coffe_batch_destroy.argtypes = [vp]
    L.coffe_batch_destroy.restype = None
    L.coffe_batch_input_size.argtypes = [vp]
    L.coffe_batch_input_size.restype = i64
    L.coffe_batch_forward.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.coffe_batch_vjp.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.coffe_batch_tables.argtypes = [vp, vp, vp, vp, vp]
    L.coffe_akima_eval.argtypes = [sz, dp, dp, sz, dp, dp]
    L.coffe_mu_weights.argtypes = [vp, i, dp, dp, dp, dp, dp, dp, dp, i, dp, dp, dp, i, i, i, i, i, dp, i, ip, i, ip,
                                   ip, i, dp, i, vp, vp]
    L.coffe_mu_weights_terms.argtypes = [vp, i, dp, dp, dp, dp, dp, dp, dp, i, dp, dp, dp, ctypes.c_uint, i, i, dp, i,
                                         ip, i, ip, ip, i, dp, i, vp, vp]
    L.coffe_mu_weights_ex.argtypes = [vp, vp, ctypes.c_uint, ctypes.c_uint, i, i, i, dp, i, ip, i, ip, ip, i, dp, i,
                                      vp, vp, vp]
    L.coffe_mu_weights_rule.argtypes = [vp, vp, ctypes.c_uint, ctypes.c_uint, i, i, i, dp, i, ip, i, ip, ip, i, dp,
                                        i, dp, dp, vp, vp, vp]
    L.coffe_mu_weights_jvp.argtypes = [vp, vp, vp, ctypes.c_uint, i, i, i, dp, i, ip, i, ip, ip, i, dp, i, vp, vp, vp]
    L.coffe_integral_table.argtypes = [i, i, sz, dp, dp, sz, d, d, dp, dp]
    L.coffe_integral_table_state.argtypes = [i, i, i, i, sz, dp, dp, sz, d, d, dp, dp]
    L.coffe_fft_any.argtypes = [dp, dp, ctypes.c_int64]
    i64ptr = ctypes.POINTER(ctypes.c_int64)
    L.coffe_flat_lensing_multipoles.argtypes = [i, dp, i, ip, i, dp, i, dp, dp, dp, ctypes.c_int64, dp, dp, i64ptr, ip,
                                                i64ptr, ip, d, d, i, i, i, dp]
    L.coffe_flat_lensing_corrfunc.argtypes = [i, dp, i, dp, i, dp, i, dp, dp, dp, i, dp, dp, d, d, i, i, i, dp]
    L.coffe_integrated_corrfunc.argtypes = [vp, ctypes.c_uint, i, ip, ip, ip, dp, dp, i, dp, dp, i, dp, i, dp, i, dp,
                                            ip, dp]
    L.coffe_divergent_tables.argtypes = [dp, dp, sz, d, d, dp, sz, dp, sz, dp, dp, dp, ctypes.POINTER(i),
                                         ctypes.POINTER(i)]
    L.coffe_batch_set_radix.argtypes = [vp, i, i, ctypes.POINTER(i)]
    L.coffe_divergent_tables_device.argtypes = [vp, vp, vp, sz, d, d, vp, sz, vp, sz, vp, vp, vp, vp, vp, vp]
    L.coffe_batch_small_r.argtypes = [vp, vp, vp, vp, vp]
    L.coffe_cov_plan_create.argtypes = [vp, i, d, d, i, ip, i, dp, ctypes.POINTER(vp)]
    L.coffe_cov_plan_destroy.argtypes = [vp]
    L.coffe_cov_plan_destroy.restype = None
    L.coffe_cov_plan_rows.argtypes = [vp]
    L.coffe_cov_plan_pairs.argtypes = [vp]
    L.coffe_cov_integrals.argtypes = [vp, i, i64p, vp, vp, vp, vp]
    L.coffe_cov_integrals_vjp.argtypes = [vp, i, i64p, vp, vp, vp, vp, vp]
    L.coffe_cov_assemble_vjp.argtypes = [vp, i, i, ip, dp, dp, dp, dp, dp, dp, i, i, vp, vp, vp, dp, dp, vp]
    L.coffe_cov_assemble.argtypes = [vp, i, i, ip, dp, dp, dp, dp, dp, dp, dp, i, i, i, i, vp, vp, vp]
    L.two_sph_bessel_status.argtypes = [dp, dp, ctypes.POINTER(dp), ctypes.c_long, ctypes.c_long, vp, dp, dp,
                                        ctypes.POINTER(dp)]
    L.coffe_terms_of.restype = ctypes.c_uint
    L.coffe_terms_of.argtypes = [i, i, i]
    L.coffe_batch_prepare.argtypes = [vp, vp, vp, vp]
    L.coffe_batch_tables_stage.argtypes = [vp, vp, vp, vp, vp]
    L.coffe_batch_project.argtypes = [vp, vp, vp, vp, vp]
    L.coffe_batch_prepare_adj.argtypes = [vp, vp, vp, vp, vp]
    L.coffe_batch_resample.argtypes = [vp, vp, vp, vp]
    L.coffe_batch_fft_prepare.argtypes = [vp, vp]
    L.coffe_batch_adj_c2r.argtypes = [vp, vp]
    L.coffe_batch_spline_adj.argtypes = [vp, vp, vp, vp]
    _lib = L
    return L


EXPORTED = [
    "coffe_engine_strerror", "coffe_engine_last_error", "coffe_engine_create", "coffe_engine_destroy",
    "coffe_engine_default", "coffe_engine_synchronize", "twofast_1bessel", "twofast_1bessel_status",
    "twofast_1bessel_noalloc", "coffe_engine_inject_alloc_failure",
    "coffe_batch_create", "coffe_batch_destroy", "coffe_batch_input_size", "coffe_batch_forward",
    "coffe_batch_vjp", "coffe_batch_tables", "coffe_batch_prepare", "coffe_batch_tables_stage",
    "coffe_batch_project", "coffe_batch_prepare_adj", "coffe_akima_eval", "coffe_background_compute", "coffe_background_jvp",
    "coffe_batch_resample", "coffe_batch_fft_prepare", "coffe_batch_adj_c2r", "coffe_batch_spline_adj",
    "coffe_mu_weights", "coffe_mu_weights_terms", "coffe_mu_weights_ex", "coffe_terms_of", "coffe_integral_table", "coffe_integral_table_state", "coffe_divergent_tables", "coffe_divergent_tables_device", "coffe_batch_set_radix", "coffe_batch_small_r",
    "coffe_cov_plan_create", "coffe_cov_plan_destroy", "coffe_cov_plan_rows", "coffe_cov_plan_pairs",
    "coffe_cov_integrals", "coffe_cov_integrals_vjp", "coffe_cov_assemble", "coffe_cov_assemble_vjp", "two_sph_bessel", "two_sph_bessel_status",
    "coffe_mu_weights_rule", "coffe_mu_weights_jvp", "coffe_fft_any", "coffe_flat_lensing_multipoles", "coffe_flat_lensing_corrfunc", "coffe_integrated_corrfunc", "coffe_refabi_set_spline_ops", "coffe_find_integral", "coffe_integrals_init",
    "coffe_integrals_free", "coffe_integrate", "coffe_multipoles_init", "coffe_multipoles_free",
    "coffe_average_multipoles_init", "coffe_average_multipoles_free", "coffe_covariance_init",
    "coffe_covariance_free", "coffe_average_covariance_free", "coffe_covariance_find", "covariance_coefficient",
]


def declare_background(L):
    dp = ctypes.POINTER(ctypes.c_double)
    d, i = ctypes.c_double, ctypes.c_int
    L.coffe_background_compute.argtypes = [d, d, d, d, d, d, i, i, dp, dp, dp, dp, dp] + [dp] * 10
    L.coffe_background_jvp.argtypes = [dp, dp, i, i, dp, dp, dp, dp, dp, dp, dp]
    return L
