/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C (C99, OpenMP) restatement of the reference algorithms on COFFE's
 * spherical-Bessel integral path: FFTLog (twofast.c), the I_l^n tables
 * (integrals.c), the local multipole quadrature (functions.c / signal.c /
 * multipoles.c) and the 2D FFTLog covariance (twobessel.c / utils.c /
 * covariance.c).  It is the parity checker for the GPU engine and the CPU
 * baseline that bench.py times.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never does.
 *
 * Pinning (tests/test_oracle_*.py): FFTs against numpy.fft and a long-double
 * DFT; splines against scipy and closed forms; Lanczos lgamma against mpmath;
 * FFTLog against closed-form Gaussian and power-law transforms and COFFE's
 * benchmark_integral fixtures; quadrature against exact polynomial integrals;
 * multipoles against the flat-sky closed form and COFFE's multipole fixtures;
 * 2D FFTLog against the closed-form Gaussian double-Bessel integral.
 */
#ifndef COFFE_ORACLE_H
#define COFFE_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_OK = 0, ORC_EINVAL = 1, ORC_ENOMEM = 2, ORC_EDOMAIN = 3, ORC_ENONFINITE = 4 };

/* FFT (orc_fft.c) */
typedef struct orc_rfft orc_rfft;
int orc_fft(size_t n, const double *in, double *out, int sign);
orc_rfft *orc_rfft_new(size_t n);
void orc_rfft_free(orc_rfft *r);
void orc_rfft_forward(orc_rfft *r, const double *x, double *X);
void orc_rfft_backward(orc_rfft *r, const double *X, double *x);
int orc_rfft_r2c(size_t n, const double *x, double *X);
int orc_rfft_c2r(size_t n, const double *X, double *x);
int orc_rfft2_r2c(size_t n1, size_t n2, const double *x, double *X);
int orc_rfft2_c2r(size_t n1, size_t n2, const double *X, double *x);

/* interpolation (orc_interp.c) */
size_t orc_bsearch(const double *xa, double x, size_t lo, size_t hi);
int orc_cspline_init(size_t n, const double *x, const double *y, double *c);
double orc_cspline_eval(size_t n, const double *x, const double *y, const double *c, double xq);
double orc_cspline_deriv(size_t n, const double *x, const double *y, const double *c, double xq);
int orc_cspline_interp(size_t n, const double *x, const double *y, size_t nq, const double *xq, double *out);
int orc_akima_init(size_t n, const double *x, const double *y, double *b, double *c, double *d);
double orc_akima_eval(size_t n, const double *x, const double *y, const double *b, const double *c,
                      const double *d, double xq);
int orc_akima_interp(size_t n, const double *x, const double *y, size_t nq, const double *xq, double *out);
typedef struct orc_bicubic orc_bicubic;
orc_bicubic *orc_bicubic_new(size_t nx, size_t ny, const double *x, const double *y, const double *z);
void orc_bicubic_free(orc_bicubic *s);
double orc_bicubic_eval(const orc_bicubic *s, double xq, double yq);
int orc_bicubic_interp(size_t nx, size_t ny, const double *x, const double *y, const double *z, size_t nq,
                       const double *xq, const double *yq, double *out);

/* quadrature and special functions (orc_quad.c) */
typedef double (*orc_fn)(double, void *);
void orc_qk61(orc_fn f, void *p, double a, double b, double *result, double *abserr, double *resabs,
              double *resasc);
double orc_qag(orc_fn f, void *p, double a, double b, double epsabs, double epsrel, size_t limit,
               double *abserr, int *status);
int orc_gauss_legendre(size_t n, double *x, double *w);
double orc_legendre_pl(int l, double x);
double orc_sph_jl(int l, double x);
double orc_qag_test(int kind, double a, double b, double epsrel, int *status);

/* FFTLog and tables (orc_fftlog.c) */
extern const double ORC_MIN_SEP[42];
int orc_twofast_window(double value, double xmin, double xmax, double xleft, double xright, double *out);
void orc_twofast_lgamma(double re, double im, double *out);
int orc_twofast_mql(double t, double q, double l, double alpha, double *out);
double orc_twofast_selectqnu(double l, double nu);
int orc_twofast_1bessel(double *output_x, double *output_y, size_t N, const double *input_x,
                        const double *input_y, size_t input_len, double l, double nu, double r0, double k0,
                        double kmin, double kmax);
int orc_integral_table(int n, int l, size_t N, size_t nin, const double *kin, const double *pin, double kmin,
                       double kmax, int with_small_r, double *out_x, double *out_y);

int orc_divergent_tables(size_t nin, const double *kin, const double *pin, double kmin, double kmax, size_t nsep,
                         const double *sep, size_t nchi, const double *chi, double *i40, double *r0, double *ren);

int orc_tables_at_sep_jvp(int nspec, const long long *off, const double *k, const double *p, const double *v,
                          const long long *nfft, int ntab, const int *tab_n, const int *tab_l, size_t nsep,
                          const double *sep, double *out, double *dout);
int orc_akima_jvp(size_t n, const double *x, const double *y, const double *dy, size_t nq, const double *xq,
                  double *out, double *dout);
int orc_tables_at_sep_batch(int nspec, const long long *off, const double *k, const double *p, const long long *nfft,
                            int ntab, const int *tab_n, const int *tab_l, size_t nsep, const double *sep, double *out);

/* multipoles (orc_multipoles.c) */
/* the d2 and g1-g3 contributions (functions.c:574-669, 856-903, 1144-1428) */
typedef struct orc_dg {
    int d2, g1, g2, g3;
    double H0h;          /* COFFE_H0 h in 1/Mpc */
    double Omega_m;      /* Omega0_cdm + Omega0_baryon */
    const double *bg_a;  /* scale factor on the background z grid */
    size_t nmb;          /* magnification and evolution biases on mb_z (Akima) */
    const double *mb_z, *s1, *s2, *fevo1, *fevo2;
    size_t ndiv;         /* the divergent table r^4 I_0^4 (Akima) */
    const double *div_x, *div_y;
    size_t nren;         /* renormalisation grid on ren_chi: ren[b nren + a] at (chi_a, chi_b) (bicubic) */
    const double *ren_chi, *ren;
    int only_cross;      /* only_cross_correlations: the d2-d2, g1-g1, g2-g2, g3-g3 terms left out */
} orc_dg;
double orc_functions_nonintegrated(const void *ctx, double z_mean, double mu, double sep);
int orc_multipoles(const double *const *tab_x, const double *const *tab_y, const size_t *tab_n, size_t nbg,
                   const double *bg_z, const double *bg_chi, const double *bg_D1, const double *bg_f,
                   const double *bg_cH, const double *bg_G1, const double *bg_G2, size_t nb, const double *zb,
                   const double *bias1, const double *bias2, int den, int rsd, int d1, int flatsky, size_t nz,
                   const double *z_mean, size_t nl, const int *ls, size_t nsep, const double *sep, int quad,
                   double epsrel, double *out);

int orc_mu_weights_hp(size_t nbg, const double *bg_z, const double *bg_chi, const double *bg_D1, const double *bg_f,
                      const double *bg_cH, const double *bg_G1, const double *bg_G2, size_t nb, const double *zb,
                      const double *bias1, const double *bias2, int den, int rsd, int d1, int flatsky, size_t nz,
                      const double *z_mean, size_t nl, const int *ls, size_t nsep, const double *sep, size_t nmu,
                      const double *mu, const double *wmu, double *out);

int orc_multipoles_dg(const double *const *tab_x, const double *const *tab_y, const size_t *tab_n, size_t nbg,
                      const double *bg_z, const double *bg_chi, const double *bg_D1, const double *bg_f,
                      const double *bg_cH, const double *bg_G1, const double *bg_G2, size_t nb, const double *zb,
                      const double *bias1, const double *bias2, int den, int rsd, int d1, int flatsky, size_t nz,
                      const double *z_mean, size_t nl, const int *ls, size_t nsep, const double *sep, int quad,
                      double epsrel, const orc_dg *dg, double *out);
/* the local correlation function at every (z_mean, mu, sep): out[z][mu][r] */
int orc_corrfunc_dg(const double *const *tab_x, const double *const *tab_y, const size_t *tab_n, size_t nbg,
                    const double *bg_z, const double *bg_chi, const double *bg_D1, const double *bg_f,
                    const double *bg_cH, const double *bg_G1, const double *bg_G2, size_t nb, const double *zb,
                    const double *bias1, const double *bias2, int den, int rsd, int d1, int flatsky, size_t nz,
                    const double *z_mean, size_t nmu, const double *mu, size_t nsep, const double *sep,
                    const orc_dg *dg, double *out);

int orc_mu_weights_hp_dg(size_t nbg, const double *bg_z, const double *bg_chi, const double *bg_D1,
                         const double *bg_f, const double *bg_cH, const double *bg_G1, const double *bg_G2, size_t nb,
                         const double *zb, const double *bias1, const double *bias2, int den, int rsd, int d1,
                         int flatsky, size_t nz, const double *z_mean, size_t nl, const int *ls, size_t nsep,
                         const double *sep, size_t nmu, const double *mu, const double *wmu, const orc_dg *dg,
                         double *out, double *out_ren);

int orc_set_threads(int n);
int orc_reference_batch(int nspec, const long long *off, const double *k, const double *p, const long long *nfft,
                        size_t nbg, const double *bg_z, const double *bg_chi, const double *bg_D1, const double *bg_f,
                        const double *bg_cH, const double *bg_G1, const double *bg_G2, size_t nb, const double *zb,
                        const double *bias1, const double *bias2, int den, int rsd, int d1, int flatsky, size_t nz,
                        const double *z_mean, size_t nl, const int *ls, size_t nsep, const double *sep, int small_r,
                        double epsrel, double *out);

/* 2D FFTLog and covariance (orc_covariance.c) */
void orc_g_l(double l, double nu, size_t n, const double *eta, double *gl);
void orc_window_2d(double *out, double width, long halfN1, long halfN2);
int orc_two_sph_bessel(const double *k1, const double *k2, const double *fk, long N1, long N2, double l1, double l2,
                       double nu1, double nu2, double width, double *r1, double *r2, double *result);
int orc_covariance_fftlog_integral(size_t nin, const double *kin, const double *pin, double power, int l1, int l2,
                                   size_t N, size_t R, const double *sep, double kmin, double kmax, double *result);
double orc_threej000(int a, int b, int c);
double orc_legendre_integral(int n, int m, int a, int b);
double orc_covariance_coefficient(double b1, double b2, double f, int ell);
int orc_covariance_assemble(size_t nz, size_t nl, const int *ls, size_t R, const double *sep, const int *pops,
                            const double *bias, const double *f, const double *D1, const double *nbar1,
                            const double *nbar2, const double *vol, const double *pix, int poisson, int mixed,
                            int cosmic, int window, const double *pk_int, const double *pk2_int, double *out);

#ifdef __cplusplus
}
#endif
#endif
