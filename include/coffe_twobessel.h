/*
 * 2D FFTLog (twobessel.h) and the Gaussian covariance of multipoles on the
 * MI355X engine.  C-ABI: plain pointers and sizes; device pointers where named d_*.
 */
#ifndef COFFE_TWOBESSEL_H
#define COFFE_TWOBESSEL_H

#include <stddef.h>
#include <stdint.h>

#include "coffe_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

/* same layout as the reference's config (twobessel.h:1-6) */
typedef struct config {
    double l1, l2;
    double nu1, nu2;
    double c_window_width;
    long Nk_sample;
} config;

/*
 * Drop-in for two_sph_bessel (twobessel.h:7, twobessel.c:16-113): r1[i] = 1/k1[N-1-i],
 * r2 likewise, result[i][j] at (r1[i], r2[j]).  The engine computes the transform of
 * a DIAGONAL fk1k2 on a square grid (k1 == k2), the form covariance_integrate_fftlog
 * calls (covariance.c:391-420); other inputs record COFFE_EUNSUPPORTED.  Odd lengths
 * (where the reference calls exit) record COFFE_EINVAL.  Even N up to 2^20: N <= 8192
 * with N and N/2 of prime factors 2, 3, 5 runs the fused LDS row kernel, other lengths
 * the general device FFT (LDS lines, four-step, Bluestein).  The void form records the
 * status for coffe_engine_last_error(); the _status form returns it.
 */
void two_sph_bessel(double *k1, double *k2, double **fk1k2, long N1, long N2, config *config, double *r1, double *r2,
                    double **result);
int two_sph_bessel_status(const double *k1, const double *k2, double **fk1k2, long N1, long N2, const config *cfg,
                          double *r1, double *r2, double **result);

/*
 * Covariance plan: N sampling points (covariance_integration_bins, even, up to 2^20;
 * lengths outside the fused LDS kernels take the general device FFT, and their
 * coffe_cov_integrals_vjp returns COFFE_EUNSUPPORTED), k_i = kmin
 * (kmax/kmin)^(i/N) with kmin/kmax the spectrum's k_min_norm/k_max_norm, multipoles
 * ls[nl], separations sep[R] (the bicubic queries).  Spectrum-independent tables are
 * built once (covariance.c:357-470 with nu1 = nu2 = 1.01, window width 0.25).
 */
typedef struct coffe_cov_plan coffe_cov_plan;
int coffe_cov_plan_create(coffe_engine *e, int N, double kmin, double kmax, int nl, const int *ls, int R,
                          const double *sep, coffe_cov_plan **out);
void coffe_cov_plan_destroy(coffe_cov_plan *p);
/* transform rows one call computes per (spectrum, power, pair); pairs l1 <= l2 */
int coffe_cov_plan_rows(const coffe_cov_plan *p);
int coffe_cov_plan_pairs(const coffe_cov_plan *p);

/*
 * The scaled FFTLog integrals of P and P^2 for nspec spectra (concatenated device
 * d_k/d_pk, host lengths n_in): d_out[s][power - 1][pair][R n + m] at (sep_m, sep_n),
 * pairs (i <= j) in row-major order over ls, times (2 l1 + 1)(2 l2 + 1) {2, 1} / pi^2
 * (covariance.c:728-767).  P is the Akima spline of each spectrum; [kmin, kmax] must
 * lie inside every spectrum's k range (else COFFE_EDOMAIN).  Synchronises the stream.
 */
int coffe_cov_integrals(coffe_cov_plan *p, int nspec, const int64_t *n_in, const double *d_k, const double *d_pk,
                        double *d_out, void *stream);

/*
 * Vector-Jacobian product of coffe_cov_integrals (no reference counterpart; the
 * differentiation of the covariance with respect to P(k)): given the cotangent
 * d_gint of d_out (same layout), d_gpk (same layout as d_pk) = d<gint, out>/dP for
 * every input sample, through the Akima spline, the diagonal k^3 P^p / dlnk and the
 * row transforms and functionals transposed.  Synchronises the stream.
 */
int coffe_cov_integrals_vjp(coffe_cov_plan *p, int nspec, const int64_t *n_in, const double *d_k, const double *d_pk,
                            const double *d_gint, double *d_gpk, void *stream);

/*
 * Gaussian covariance of the multipoles (covariance.c:790-1060, the FFTLog branch), per
 * z bin k: the galaxy biases 1..4 at bias[4 k .. 4 k + 3], growth rate f[k], D1[k], the
 * number densities of the first and second tracer nbar1[k] and nbar2[k] (density1 /
 * density2), volume vol[k], pixel size pix[k] (host arrays).  pops[0..3] are the
 * populations of the four fields (covariance_pop1..4, each 1..4; NULL means 1, 1, 1, 1,
 * the reference's default, parser.c:945-948): a coefficient of populations (a, b) reads
 * biases bias[a - 1] and bias[b - 1] (covariance.c:60-74) and the mixed and Poisson terms
 * carry the reference's Kronecker deltas of the populations with density1 / density2
 * (covariance.c:889-967).  poisson/mixed/cosmic select the terms; window selects the
 * bin-averaged Poisson term (covariance_window, covariance.c:931-944).  d_integ from
 * coffe_cov_integrals; d_out[s][z][l1][r1][l2][r2] (device).  COFFE_EINVAL for a
 * population outside 1..4, COFFE_EDOMAIN for a non-positive density, volume or pixel.
 */
int coffe_cov_assemble(coffe_cov_plan *p, int nspec, int nz, const int *pops, const double *bias, const double *f,
                       const double *D1, const double *nbar1, const double *nbar2, const double *vol,
                       const double *pix, int poisson, int mixed, int cosmic, int window, const double *d_integ,
                       double *d_out, void *stream);

/*
 * Vector-Jacobian product of coffe_cov_assemble (no reference counterpart; the
 * differentiation of the covariance): given the cotangent d_gout of d_out, writes
 * d_ginteg (layout of d_integ) = d<gout, cov>/d integrals, and on the host
 * out_gbias[s][z][4] and out_gf[s][z] = d<gout_s, cov_s>/d galaxy bias 1..4 and /d f of
 * every z bin, per spectrum (either may be NULL).  The Poisson term depends on neither.
 * Same populations and per-bin arguments as coffe_cov_assemble; synchronises the stream.
 */
int coffe_cov_assemble_vjp(coffe_cov_plan *p, int nspec, int nz, const int *pops, const double *bias,
                           const double *f, const double *D1, const double *nbar1, const double *nbar2,
                           const double *vol, int mixed, int cosmic, const double *d_integ, const double *d_gout,
                           double *d_ginteg, double *out_gbias, double *out_gf, void *stream);

#ifdef __cplusplus
}
#endif
#endif
