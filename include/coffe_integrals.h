/*
 * Drop-in for the reference's table entry point (integrals.h, integrals.c:459-573),
 * for callers outside the reference tree; the reference's own callers keep
 * integrals.h, whose declaration this one matches in ABI (the two enum parameters
 * are passed as int, their underlying type).  Do not include both.
 */
#ifndef COFFE_INTEGRALS_DROPIN_H
#define COFFE_INTEGRALS_DROPIN_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* the reference's coffe_interpolation (common.h:228-232): a gsl_spline * (GSL's
   gsl_spline: interp, x, y, size) and a gsl_interp_accel *; only the spline's
   x, y and size are read */
struct coffe_interpolation;

/*
 * coffe_integrals_renormalizable (integrals.c:459-573): *output_x / *output_y are
 * malloc'ed here (the caller frees them) with *real_output_len = output_len + 43
 * entries: r = 0, the 42 min_sep values, the output_len FFTLog nodes; y = I_l^n(r),
 * times r^(n-l) when n >= l.  The FFTLog runs on the spectrum's spline samples
 * (twofast_1bessel with nu = n, r0 = 1, k0 = kmin = x_min, kmax = x_max), the
 * small-r nodes are QAG-GK61 integrals over [x_min, x_max] of their Akima spline
 * (the reference's default spline type).  Both states COFFE_INTEGER (0) or both
 * COFFE_HALF_INTEGER (1, the flat-sky lensing tables: coffe_integral_table_state);
 * mixed states return COFFE_EUNSUPPORTED.  Returns 0 or a COFFE_E* code (the reference exits instead);
 * on error nothing is allocated.
 */
int coffe_integrals_renormalizable(double **output_x, double **output_y, const size_t output_len,
                                   size_t *real_output_len, const struct coffe_interpolation *spectrum, const int n,
                                   const int l, const int state_n, const int state_l, const double x_min,
                                   const double x_max);

#ifdef __cplusplus
}
#endif
#endif
