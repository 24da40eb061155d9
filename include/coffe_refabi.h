/*
 * The reference's C entry points of the table, multipole and covariance paths, for a
 * COFFE C build that links this library in place of integrals.c, signal.c,
 * multipoles.c, average_multipoles.c and covariance.c.
 *
 * The structures below restate the reference's layouts field by field (common.h:
 * 86-333 and 340-550, background.h:23-44, integrals.h:22-52, multipoles.h:27-41,
 * average_multipoles.h:23-34, covariance.h:24-46) for the C build of COFFE
 * (COFFE_CYTHON undefined, so coffe_parameters_t holds the libconfig pointer `conf`),
 * so that the reference's callers (main.c, output.c) pass their own objects unchanged.
 * The reference's own headers declare the same functions with the same ABI: a COFFE
 * build includes those, and this header is for callers outside the reference tree.
 * Do not include both.
 *
 * The reference's interpolation objects (GSL splines behind coffe_interpolation) are
 * made, evaluated and freed by the host program's own coffe_init_spline,
 * coffe_interp_spline, coffe_init_spline2d, coffe_free_spline and coffe_free_spline2d
 * (common.c), which a COFFE build links in; this library refers to them weakly.  A
 * host without them registers equivalents with coffe_refabi_set_spline_ops.  The node
 * arrays of the power spectrum, background and bias splines (GSL's gsl_spline: interp,
 * x, y, size; gsl_spline2d: the gsl_interp2d object, then xarr, yarr, zarr) are read
 * directly.
 *
 * Every function returns 0 (EXIT_SUCCESS) or a COFFE_E* code of coffe_engine.h where the
 * reference exits or has no error path; coffe_engine_last_error() also returns that code.
 */
#ifndef COFFE_REFABI_H
#define COFFE_REFABI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define COFFE_REFABI_MAX_STRLEN 256   /* COFFE_MAX_STRLEN, common.h:47-49 */

/* common.h:228-245 */
typedef struct coffe_interpolation {
    void *spline;   /* gsl_spline * */
    void *accel;    /* gsl_interp_accel * */
} coffe_interpolation;

typedef struct coffe_interpolation2d {
    void *spline;   /* gsl_spline2d * */
    void *xaccel, *yaccel;
} coffe_interpolation2d;

/* enums of common.h:251-317 (passed as int, their underlying type) */
enum { COFFE_REF_NONINTEGRATED = 0, COFFE_REF_SINGLE_INTEGRATED = 1, COFFE_REF_DOUBLE_INTEGRATED = 2 };
enum {
    COFFE_REF_CORRFUNC = 1,
    COFFE_REF_MULTIPOLES = 2,
    COFFE_REF_AVERAGE_MULTIPOLES = 3,
    COFFE_REF_COVARIANCE_MULTIPOLES = 4,
    COFFE_REF_COVARIANCE_AVERAGE_MULTIPOLES = 5
};
enum { COFFE_REF_INTERP_CSPLINE = 3, COFFE_REF_INTERP_AKIMA = 5 };
enum { COFFE_REF_INTERP2D_BILINEAR = 1, COFFE_REF_INTERP2D_BICUBIC = 2 };
enum { COFFE_REF_COVARIANCE_STANDARD = 1, COFFE_REF_COVARIANCE_FFTLOG = 2 };
enum { COFFE_REF_INTEGRATION_GSL = 1, COFFE_REF_INTEGRATION_DOUBLE_EXPONENTIAL = 2 };
enum { COFFE_REF_INTEGER = 0, COFFE_REF_HALF_INTEGER = 1 };   /* integrals.h:23 */

/* common.h:329-333 */
typedef struct coffe_correlation_contributions {
    int den, rsd, len, d1, d2, g1, g2, g3, g4, g5;
} coffe_correlation_contributions;

/* common.h:86-117 */
typedef struct coffe_class_struct_t {
    void *file_content, *background, *thermodynamics, *perturb, *primordial, *nonlinear, *transfer, *spectra;
} coffe_class_struct_t;

typedef struct coffe_class_precision_parameters_t {
    double k_per_decade_for_bao, k_per_decade_for_pk, start_large_k_at_tau_h_over_tau_k;
    int l_max_g, l_max_ur;
    double tol_perturb_integration, radiation_streaming_trigger_tau_over_tau_k, ur_fluid_trigger_tau_over_tau_k;
} coffe_class_precision_parameters_t;

/* common.h:150-153 */
typedef struct coffe_fit_coefficients_array_t {
    void *array;
    size_t size;
} coffe_fit_coefficients_array_t;

/* common.h:340-550, in the reference's order */
typedef struct coffe_parameters_t {
    int output_type;
    double *mu;
    size_t mu_len;
    coffe_correlation_contributions correlation_contrib;
    char **type_bg;
    size_t type_bg_len;
    int background_bins, bessel_bins;
    double inv_k_window, Omega0_cdm, Omega0_m;
    int b_derivative, f_derivative, b_tilde_derivative, f_tilde_derivative;
    double Omega0_baryon, Omega0_gamma, w0, wa, Omega0_de;
    double *z_mean;
    size_t z_mean_len;
    double *deltaz;
    size_t deltaz_len;
    double *sep;
    size_t sep_len;
    int integration_method, integration_bins, nthreads;
    char file_power_spectrum[COFFE_REFABI_MAX_STRLEN];
    coffe_interpolation power_spectrum, power_spectrum_norm;
    coffe_interpolation2d power_spectrum2d, power_spectrum2d_norm;
    double k_min, k_max, k_min_norm, k_max_norm;
    int read_galaxy_bias1, read_galaxy_bias2;
    char file_galaxy_bias1[COFFE_REFABI_MAX_STRLEN], file_galaxy_bias2[COFFE_REFABI_MAX_STRLEN];
    coffe_interpolation galaxy_bias1, galaxy_bias2, galaxy_bias3, galaxy_bias4;
    int read_magnification_bias1, read_magnification_bias2;
    int galaxy_bias_analytic;
    char file_magnification_bias1[COFFE_REFABI_MAX_STRLEN], file_magnification_bias2[COFFE_REFABI_MAX_STRLEN];
    coffe_interpolation magnification_bias1, magnification_bias2;
    int read_evolution_bias1, read_evolution_bias2;
    char file_evolution_bias1[COFFE_REFABI_MAX_STRLEN], file_evolution_bias2[COFFE_REFABI_MAX_STRLEN];
    coffe_interpolation evolution_bias1, evolution_bias2;
    int divergent;
    void *conf;   /* config_t *, present in the C build */
    char timestamp[COFFE_REFABI_MAX_STRLEN], output_path[COFFE_REFABI_MAX_STRLEN],
        output_prefix[COFFE_REFABI_MAX_STRLEN];
    int interp_method;
    int *multipole_values;
    size_t multipole_values_len;
    int has_class, has_cuba;
    double *zmin;
    size_t zmin_len;
    double *zmax;
    size_t zmax_len;
    double *fsky;
    size_t fsky_len;
    double *density1, *density2;
    size_t density1_len, density2_len;
    double *pixelsize;
    size_t pixelsize_len;
    int covariance_pop1, covariance_pop2, covariance_pop3, covariance_pop4;
    int covariance_cosmic, covariance_mixed, covariance_poisson;
    int covariance_integration_method, covariance_integration_bins;
    int covariance_interpolation_method;
    int covariance_window, have_window;
    double window_size;
    int flatsky_local, flatsky_local_nonlocal, flatsky_nonlocal, verbose;
    int have_class;
    double n_s, sigma8, A_s, T_cmb, N_ur, m_ncdm, Omega0_nu, YHe;
    int N_ncdm;
    double h, k_pivot;
    coffe_class_struct_t class_struct;
    coffe_class_precision_parameters_t class_precision;
    int flag;
    int pk_type;
    int midpoint_approximation, only_cross_correlations;
    int integration_1d_type;
    double integration_1d_prec;
    coffe_fit_coefficients_array_t galaxy_bias1_coefficients, galaxy_bias2_coefficients,
        magnification_bias1_coefficients, magnification_bias2_coefficients, evolution_bias1_coefficients,
        evolution_bias2_coefficients;
    int degree_galaxy_bias1, degree_galaxy_bias2, degree_magnification_bias1, degree_magnification_bias2,
        degree_evolution_bias1, degree_evolution_bias2;
} coffe_parameters_t;

/* background.h:23-44: every spline on the same z nodes, except z_as_chi (on chi) */
typedef struct coffe_background_t {
    coffe_interpolation z_as_chi, a, Hz, conformal_Hz, conformal_Hz_prime, D1, f, G1, G2, comoving_distance;
    int flag;
} coffe_background_t;

/* integrals.h:25-52 */
typedef struct coffe_integral_t {
    coffe_interpolation result;
    coffe_interpolation2d result2d, renormalization;
    coffe_interpolation renormalization_zero_separation;
    int n, l;
    int state_n, state_l;
} coffe_integral_t;

typedef struct coffe_integral_array_t {
    coffe_integral_t *array;
    size_t size;
} coffe_integral_array_t;

/* common.h:178-217 and multipoles.h / average_multipoles.h / covariance.h */
typedef struct coffe_multipoles_coords_t {
    double z_mean, separation;
    int l;
} coffe_multipoles_coords_t;
typedef struct coffe_multipoles_t {
    coffe_multipoles_coords_t coords;
    double value;
} coffe_multipoles_t;
typedef struct coffe_multipoles_array_t {
    coffe_multipoles_t *array;
    size_t size;
} coffe_multipoles_array_t;

typedef struct coffe_average_multipoles_coords_t {
    double z_min, z_max, separation;
    int l;
} coffe_average_multipoles_coords_t;
typedef struct coffe_average_multipoles_t {
    coffe_average_multipoles_coords_t coords;
    double value;
} coffe_average_multipoles_t;
typedef struct coffe_average_multipoles_array_t {
    coffe_average_multipoles_t *array;
    size_t size;
} coffe_average_multipoles_array_t;

typedef struct coffe_covariance_coords_t {
    int l1, l2;
    double separation1, separation2, z_mean;
} coffe_covariance_coords_t;
typedef struct coffe_covariance_t {
    coffe_covariance_coords_t coords;
    double value;
} coffe_covariance_t;
typedef struct coffe_covariance_array_t {
    coffe_covariance_t *array;
    size_t size;
} coffe_covariance_array_t;

typedef struct coffe_average_covariance_coords_t {
    int l1, l2;
    double separation1, separation2, z_min, z_max;
} coffe_average_covariance_coords_t;
typedef struct coffe_average_covariance_t {
    coffe_average_covariance_coords_t coords;
    double value;
} coffe_average_covariance_t;
typedef struct coffe_average_covariance_array_t {
    coffe_average_covariance_t *array;
    size_t size;
} coffe_average_covariance_array_t;

/*
 * The host's interpolation routines (common.c: coffe_init_spline, coffe_init_spline2d,
 * coffe_interp_spline, coffe_interp_spline2d, coffe_free_spline, coffe_free_spline2d,
 * same signatures with the enums as int).  No reference counterpart: a COFFE build
 * needs no call, its own routines are found; a host without them (a test harness, a
 * program statically linked without exporting them) registers its own.  NULL
 * returns to the host's.  Not thread-safe against concurrent entry-point calls.
 */
typedef struct coffe_spline_ops {
    int (*init_spline)(coffe_interpolation *interp, const double *x, const double *y, size_t n, int type);
    int (*init_spline2d)(coffe_interpolation2d *interp, const double *x, const double *y, const double *z, size_t nx,
                         size_t ny, int type);
    double (*interp_spline)(const coffe_interpolation *interp, double x);
    double (*interp_spline2d)(const coffe_interpolation2d *interp, double x, double y);
    int (*free_spline)(coffe_interpolation *interp);
    int (*free_spline2d)(coffe_interpolation2d *interp);
} coffe_spline_ops;
void coffe_refabi_set_spline_ops(const coffe_spline_ops *ops);

/*
 * coffe_find_integral (integrals.h:56-62, integrals.c:140-169): the entry with these
 * (n, l, state_n, state_l), or NULL.
 */
coffe_integral_t *coffe_find_integral(const coffe_integral_array_t *integral, int n, int l, int state_n,
                                      int state_l);

/*
 * coffe_integrals_init (integrals.h:84-88, integrals.c:579-1424): the eight tables I_l^n
 * (n, l) = (0,0) (0,2) (0,4) (1,1) (1,3) (2,0) (2,2) (3,1) of power_spectrum_norm over
 * [k_min_norm, k_max_norm] with bessel_bins FFTLog nodes and the 43 small-separation
 * nodes (the engine's coffe_integral_table), each put in `result` with the host's
 * coffe_init_spline and interp_method; with `divergent` set also I_0^4 at the
 * reference's fixed separation list (integrals.c:76-96; result), its zero-separation
 * renormalisation (renormalization_zero_separation, first node (0, 0)) and the 201 x 201
 * grid of the renormalisation on [0, chi_max] (renormalization,
 * covariance_interpolation_method), chi_max = chi(max(z_mean + deltaz)) for CORRFUNC / MULTIPOLES output and
 * chi(max(zmax)) for AVERAGE_MULTIPOLES; other outputs build no grid (the reference's
 * grid would be degenerate there).  With flatsky_local_nonlocal or flatsky_nonlocal
 * also the flat-sky lensing tables (integrals.c:915-962, 1082-1130, 1246-1300): the
 * half-integer (1/2, -1/2) one, kept times sqrt(2/pi) 2 pi^2 as there, and for
 * MULTIPOLES output the (1, l) tables (flatsky_nonlocal) and the half-integer
 * ((l - 2k + 3)/2, (l - 2k + 1)/2) tables, k <= l/2 (flatsky_local_nonlocal), of the
 * even multipoles (the odd multipoles' terms are zero and their tables unused; tables
 * the array holds already are not repeated); lensing-lensing multipoles above l = 4
 * and CLASS spectra (pk_type != 0 with midpoint_approximation) return
 * COFFE_EUNSUPPORTED.  The array is freed first.
 */
int coffe_integrals_init(const coffe_parameters_t *par, const coffe_background_t *bg,
                         coffe_integral_array_t *integral);
/* coffe_integrals_free (integrals.h:90-92, integrals.c:1426-1444) */
int coffe_integrals_free(coffe_integral_array_t *integral);

/*
 * coffe_integrate (signal.h:17-29, signal.c:472-771) for the local terms: flag_integral
 * NONINTEGRATED with flag_output CORRFUNC (the local correlation function at (z_mean,
 * sep, mu)), MULTIPOLES (the multipole l at (z_mean, sep), a 256-node Gauss-Legendre mu
 * rule in place of the reference's adaptive one) or AVERAGE_MULTIPOLES (the multipole
 * averaged over [z_min, z_max], a 32-node Gauss-Legendre rule in redshift).  The
 * integrated terms (signal.c:556-771) for CORRFUNC and MULTIPOLES: SINGLE_INTEGRATED
 * gives the flat-sky density-lensing term (len with den and flatsky_local_nonlocal;
 * RSD- and Doppler-lensing are zero there), DOUBLE_INTEGRATED the flat-sky
 * lensing-lensing term (len with flatsky_nonlocal, not with only_cross_correlations),
 * both from the integral array's tables on the GPU (coffe_flat_lensing_multipoles /
 * coffe_flat_lensing_corrfunc); they return 0 when the reference would (no len / g4 /
 * g5 contribution) and NaN with COFFE_EUNSUPPORTED where the reference integrates by
 * Monte Carlo (full-sky lensing, g4, g5, lensing with d2 / g1-g3, AVERAGE_MULTIPOLES).
 * Errors return NaN; coffe_engine_last_error() gives the code.
 */
double coffe_integrate(coffe_parameters_t *par, coffe_background_t *bg, coffe_integral_array_t *integral,
                       double z_mean, double z_min, double z_max, double sep, double mu, int l, int flag_integral,
                       int flag_output);

/*
 * coffe_multipoles_init (multipoles.h:47-52, multipoles.c:48-136): mp->array of
 * z_mean_len x multipole_values_len x sep_len records, z-major then l then separation,
 * each the local multipole (coffe_integrate NONINTEGRATED / MULTIPOLES) plus, with len,
 * the flat-sky lensing terms (SINGLE_ and DOUBLE_INTEGRATED as coffe_integrate gives
 * them), computed in one batch on the GPU.  g4, g5 and the lensing terms that need Monte
 * Carlo return COFFE_EUNSUPPORTED.  The array is freed first.
 */
int coffe_multipoles_init(coffe_parameters_t *par, coffe_background_t *bg, coffe_integral_array_t *integral,
                          coffe_multipoles_array_t *mp);
/* multipoles.c:138-150 */
int coffe_multipoles_free(coffe_multipoles_array_t *mp);

/*
 * coffe_average_multipoles_init (average_multipoles.h:36-41, average_multipoles.c:
 * 64-146): zmin_len x multipole_values_len x sep_len records, bin-major then l then
 * separation, each the redshift-averaged local multipole (as coffe_integrate
 * AVERAGE_MULTIPOLES).
 */
int coffe_average_multipoles_init(coffe_parameters_t *par, coffe_background_t *bg, coffe_integral_array_t *integral,
                                  coffe_average_multipoles_array_t *ramp);
int coffe_average_multipoles_free(coffe_average_multipoles_array_t *ramp);

/*
 * coffe_covariance_init (covariance.h:48-53, covariance.c:477-1103), the FFTLog method
 * (covariance_integration_method 2, or 1, the STANDARD quadrature, which is computed by
 * FFTLog as well; bicubic interpolation): the Gaussian covariance of
 * the multipoles per z_mean bin (COVARIANCE_MULTIPOLES, records z, l1, r1, l2, r2) or
 * per [zmin, zmax] bin (COVARIANCE_AVERAGE_MULTIPOLES: z_mean = (zmin + zmax)/2 and the
 * volume 4 pi fsky / (int dz / (H chi^2 (1 + z)) (H0 h)^2), covariance.c:808-837), with
 * the populations, densities, pixel sizes and term switches of par.  The integrals of P
 * and P^2 run on the GPU (coffe_cov_integrals over power_spectrum_norm, k_min_norm,
 * k_max_norm, covariance_integration_bins points); the biases, f and D1 at each bin
 * come from the host's splines.  Other output types leave both arrays untouched; the
 * bilinear interpolation and CLASS spectra return COFFE_EUNSUPPORTED.
 */
int coffe_covariance_init(coffe_parameters_t *par, coffe_background_t *bg, coffe_covariance_array_t *cov_mp,
                          coffe_average_covariance_array_t *cov_ramp);
/* covariance.c:1105-1126 */
int coffe_covariance_free(coffe_covariance_array_t *cov);
int coffe_average_covariance_free(coffe_average_covariance_array_t *cov);
/*
 * coffe_covariance_find (covariance.h:65-72, covariance.c:82-110): the first record
 * with these l1, l2 and separations and z_mean within a relative 1e-3.  The reference
 * exits when none matches; here the returned record has value NaN and l1 = l2 = -1,
 * and coffe_engine_last_error() returns COFFE_EDOMAIN.
 */
coffe_covariance_t coffe_covariance_find(const coffe_covariance_array_t *cov, double z_mean, int l1, int l2,
                                         double sep1, double sep2);
/* covariance_coefficient (covariance.h:74-79, covariance.c:40-57) */
double covariance_coefficient(double bias1, double bias2, double growth_rate, int ell);

#ifdef __cplusplus
}
#endif
#endif
