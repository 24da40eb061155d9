/*
 * coffe_engine.h -- C-ABI of the MI355X spherical-Bessel integral engine.
 *
 * Two groups of entry points:
 *
 *  1. Drop-in replacements with the reference's own signatures, so that COFFE's
 *     callers link unchanged:
 *       twofast_1bessel        replaces twofast.c:362-483 (decl. twofast.h:44-58)
 *       two_sph_bessel         replaces twobessel.c:16-113 (decl. twobessel.h:7)
 *     Both are void in the reference and exit() on bad input; here they record an
 *     error code readable with coffe_engine_last_error() and never exit.  The
 *     *_status variants return the code directly.
 *
 *  2. The batched engine: many spectra with ragged k-grids and FFT lengths, any set
 *     of I_l^n tables, evaluated at a common set of separations and projected onto
 *     multipoles with spectrum-independent weights A[z][l][t][r] (the mu-integrals
 *     of functions_nonintegrated, functions.c:89-1484, see coffe_mu_weights_*),
 *     plus the vector-Jacobian product with respect to every input P(k) sample.
 *     All array arguments of the batch calls are device pointers; the stream is a
 *     hipStream_t (NULL = the engine's own stream).
 *
 * All functions return 0 on success or one of the COFFE_E* codes.  Handles are
 * re-entrant per handle; concurrent calls on one handle must be serialised by the
 * caller.
 */
#ifndef COFFE_ENGINE_H
#define COFFE_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    COFFE_OK = 0,
    COFFE_EINVAL = 1,        /* bad argument (sizes, ordering, NULL) */
    COFFE_ENOMEM = 2,        /* host or device allocation failed */
    COFFE_EDEVICE = 3,       /* HIP runtime / kernel launch error */
    COFFE_EUNSUPPORTED = 4,  /* e.g. FFT length with a prime factor > 5 */
    COFFE_EDOMAIN = 5,       /* window limits, separation outside the table */
    COFFE_ENONFINITE = 6,    /* NaN/inf in an input or in the gamma-ratio kernel */
    COFFE_ENODEVICE = 7,     /* no HIP device / library built without one */
    COFFE_ENOCONV = 8        /* adaptive quadrature did not converge (interval limit, roundoff) */
};

typedef struct coffe_engine coffe_engine;
typedef struct coffe_batch coffe_batch;

const char *coffe_engine_strerror(int code);
/* last error recorded by a void drop-in entry point on the calling thread */
int coffe_engine_last_error(void);

int coffe_engine_create(int device, coffe_engine **out);
/* testing hook (no reference counterpart): the n-th device allocation of the library
   from now on fails with COFFE_ENOMEM (n >= 1; 0 switches the injection off), so that
   tests can drive the clean-up paths of every allocating call */
void coffe_engine_inject_alloc_failure(int n);
void coffe_engine_destroy(coffe_engine *e);
/* the engine used by the drop-in entry points (created on first use, device 0) */
coffe_engine *coffe_engine_default(void);

/* ---- reference drop-ins -------------------------------------------------- */

/* twofast.c:362-483: 1D FFTLog of an input P(k) table (natural cubic spline
   resampled onto output_len log points); host buffers, flag ignored (FFTW effort).
   Any even output_len up to 2^21: lengths with N/2 = 2^a 3^b 5^c <= 8192 run the fused
   LDS kernels, the others the general device transform of coffe_fft_any. */
void twofast_1bessel(double *output_x, double *output_y, size_t output_len, double *input_x, double *input_y,
                     size_t input_len, double l, double nu, double r0, double k0, double kmin, double kmax,
                     unsigned flag);
int twofast_1bessel_status(double *output_x, double *output_y, size_t output_len, const double *input_x,
                           const double *input_y, size_t input_len, double l, double nu, double r0, double k0,
                           double kmin, double kmax);

/*
 * twofast_1bessel_status on device arrays, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  d_input_x/y
 * (input_len samples) and d_output_x/y (output_len values, output_x = r0 (kmax /
 * kmin)^(i/N)) are device arrays on the current device; all workspace, including the
 * grid tables (recomputed every call, nothing is cached), lives in one fixed device
 * array of the library (about 1.2 MB per GPU, kept until the process ends).  Limits:
 * output_len <= 16384 (even, N/2 = 2^a 3^b 5^c), input_len <= 65536
 * (COFFE_EUNSUPPORTED beyond).  The call is synchronous on `stream` (NULL = the null
 * stream); calls in one process are serialised by a lock, and every call on a GPU
 * reuses that GPU's array, so calls that overlap on one GPU from several processes
 * race.  Returns COFFE_ENONFINITE for a non-finite input sample or gamma-ratio kernel
 * (checked before the transform runs; the outputs are then not written), and the
 * validation codes of twofast_1bessel_status otherwise.
 */
int twofast_1bessel_noalloc(double *d_output_x, double *d_output_y, size_t output_len, const double *d_input_x,
                            const double *d_input_y, size_t input_len, double l, double nu, double r0, double k0,
                            double kmin, double kmax, void *stream);

/* host Akima interpolation (common.c:524-576, 704-710 with the default Akima type):
   out[q] = spline(x, y)(xq[q]), NaN outside [x_0, x_{n-1}] */
int coffe_akima_eval(size_t n, const double *x, const double *y, size_t nq, const double *xq, double *out);

/* ---- batched engine ------------------------------------------------------- */

/*
 * nspec spectra; spectrum s has n_in[s] input samples (k, P) stored contiguously
 * (spectrum after spectrum) and FFT length n_fft[s] (even, N/2 = 2^a 3^b 5^c,
 * N <= 16384).  kmin[s] / kmax[s] must equal its first / last k (as
 * set_power_spectrum_linear does, coffe.pyx:2961-2964).  Tables t = 0..ntab-1 are
 * I_{tab_l[t]}^{tab_n[t]} (stored as r^(n-l) I when n >= l, integrals.c:514-522).
 * The nsep separations (host array, Mpc) must lie strictly inside the FFT part of
 * every table's r-grid (r >= (kmax/kmin)^(2/N)); nz * nl is the number of
 * (redshift, multipole) rows of the weights A.
 */
int coffe_batch_create(coffe_engine *e, int nspec, const int64_t *n_in, const int64_t *n_fft, const double *kmin,
                       const double *kmax, int ntab, const int *tab_n, const int *tab_l, int nsep, const double *sep,
                       int nz, int nl, coffe_batch **out);
void coffe_batch_destroy(coffe_batch *b);
/* total number of input samples (sum n_in) and of output values per call */
int64_t coffe_batch_input_size(const coffe_batch *b);

/*
 * Forward: itab[s][t][r] = I_t(sep_r) for spectrum s (may be NULL) and
 * out[s][z][l][r] = sum_t A[z][l][t][r] itab[s][t][r] (may be NULL).
 */
int coffe_batch_forward(coffe_batch *b, const double *d_k, const double *d_pk, const double *d_A, double *d_itab,
                        double *d_out, void *stream);

/*
 * Forward plus vector-Jacobian product: given the cotangent gout[s][z][l][r] of
 * out, writes gpk (same layout as pk) = d<gout, out>/dP, and optionally itab/out.
 */
int coffe_batch_vjp(coffe_batch *b, const double *d_k, const double *d_pk, const double *d_A, const double *d_gout,
                    double *d_itab, double *d_out, double *d_gpk, void *stream);

/* the FFT pass order of the batch's transforms of length M = n_fft/2: radix[0..npass)
   from {2, 3, 4, 5, 8} with product M, first pass first (npass <= 10); the default is
   radix 8 first, then 4, 2, 5, 3.  Synchronous (one upload); COFFE_EDOMAIN when the
   batch has no transform of that length.  Results change at rounding level only. */
int coffe_batch_set_radix(coffe_batch *b, int M, int npass, const int *radix);

/* full tables: full[s][t][i] = r^(n-l) I_t(r_i), r_i = (kmax/kmin)^(i/N), i < n_fft[s];
   row stride = max n_fft */
int coffe_batch_tables(coffe_batch *b, const double *d_k, const double *d_pk, double *d_full, void *stream);

/*
 * The stages that coffe_batch_forward / coffe_batch_vjp launch, exposed so that a
 * caller can interleave its own work or time each kernel on its stream:
 *   prepare      spline + resample + r2c of every spectrum (needs d_k, d_pk)
 *   tables       kernel multiply + c2r + Akima at the separations; with d_gout
 *                also the Akima VJP and the transposed c2r (adjoint mode)
 *   project      out = sum_t A itab
 *   prepare_adj  transposed r2c, resample and spline -> gpk
 * Stages must run in this order on one stream; the batch keeps the workspace.
 */
int coffe_batch_prepare(coffe_batch *b, const double *d_k, const double *d_pk, void *stream);
int coffe_batch_tables_stage(coffe_batch *b, const double *d_A, const double *d_gout, double *d_itab, void *stream);
int coffe_batch_project(coffe_batch *b, const double *d_A, const double *d_itab, double *d_out, void *stream);
int coffe_batch_prepare_adj(coffe_batch *b, const double *d_k, const double *d_pk, double *d_gpk, void *stream);
/* the single kernels behind prepare (resample, then fft_prepare) and prepare_adj
   (adj_c2r, then spline_adj), same ordering rules */
int coffe_batch_resample(coffe_batch *b, const double *d_k, const double *d_pk, void *stream);
int coffe_batch_fft_prepare(coffe_batch *b, void *stream);
int coffe_batch_adj_c2r(coffe_batch *b, void *stream);
int coffe_batch_spline_adj(coffe_batch *b, const double *d_k, double *d_gpk, void *stream);

/* terms of the local correlation function (functions.c:129-1143): auto terms and
   the cross terms of density, RSD and Doppler-1 */
enum {
    COFFE_TERM_DEN_DEN = 1,
    COFFE_TERM_RSD_RSD = 2,
    COFFE_TERM_D1_D1 = 4,
    COFFE_TERM_DEN_RSD = 8,
    COFFE_TERM_DEN_D1 = 16,
    COFFE_TERM_RSD_D1 = 32,
    /* the d2 and g1-g3 terms (functions.c:574-669, 856-903, 1144-1428): among the chosen
       d2 / g sources (their auto and mutual terms, on I_0^4 minus its renormalisation),
       and their crosses with density (I_0^2), RSD (I_0^2, I_2^2) and Doppler-1 (I_1^3) */
    COFFE_TERM_DG_DG = 64,
    COFFE_TERM_DEN_DG = 128,
    COFFE_TERM_RSD_DG = 256,
    COFFE_TERM_D1_DG = 512,
    /* with COFFE_TERM_DG_DG: only the mutual terms among the d2 / g sources, their autos
       left out (only_cross_correlations, functions.c:574-669) */
    COFFE_TERM_DG_MUTUAL = 1024
};

/* the d2 / g sources of coffe_mu_weights_ex */
enum { COFFE_SOURCE_D2 = 1, COFFE_SOURCE_G1 = 2, COFFE_SOURCE_G2 = 4, COFFE_SOURCE_G3 = 8 };

/* the term mask of the reference's contribution flags: auto terms of each flag set
   and cross terms of each pair set (functions.c:129-1143) */
unsigned coffe_terms_of(int den, int rsd, int d1);

/*
 * mu-quadrature weights of the local multipoles (functions_nonintegrated,
 * functions.c:89-1484, integrated as in signal.c:83-111, 521-542):
 *   A[z][l][t][r] = (2l+1)/2 sum_q w_q P_l(mu_q) G c_t(z_mean[z], mu_q, sep[r])
 * with an nmu-node Gauss-Legendre rule, so that multipoles = sum_t A I_t(sep) (the
 * project stage).  Background tables (length nbg, on the z grid bg_z; chi is
 * also used inverted as z(chi)) and the galaxy-bias tables (nb samples on bias_z)
 * are host arrays, interpolated with Akima as the reference does.  Contributions:
 * den, rsd, d1 (0/1), flatsky selects the flat-sky forms.  Tables t are (tab_n,
 * tab_l) among the 8 defaults and must include every table the contributions use.
 * d_A is device memory [nz][nl][ntab][nsep]; nl <= 4.  Asynchronous on `stream`
 * (NULL: the engine's stream): the host arrays may be reused when the call returns,
 * and d_A is complete when the stream reaches this point.
 */
int coffe_mu_weights(coffe_engine *e, int nbg, const double *bg_z, const double *bg_chi, const double *bg_D1,
                     const double *bg_f, const double *bg_cH, const double *bg_G1, const double *bg_G2, int nb,
                     const double *bias_z, const double *bias1, const double *bias2, int den, int rsd, int d1,
                     int flatsky, int nz, const double *zmean, int nl, const int *ls, int ntab, const int *tab_n,
                     const int *tab_l, int nsep, const double *sep, int nmu, double *d_A, void *stream);

/* the same for an explicit COFFE_TERM_* mask: splitting the terms gives the weights'
   polynomial dependence on a constant bias (den-den ~ b^2, den-rsd and den-d1 ~ b)
   and on a constant growth rate (rsd-rsd, d1-d1, rsd-d1 ~ f^2; den-rsd, den-d1 ~ f) */
int coffe_mu_weights_terms(coffe_engine *e, int nbg, const double *bg_z, const double *bg_chi, const double *bg_D1,
                           const double *bg_f, const double *bg_cH, const double *bg_G1, const double *bg_G2, int nb,
                           const double *bias_z, const double *bias1, const double *bias2, unsigned terms,
                           int flatsky, int nz, const double *zmean, int nl, const int *ls, int ntab,
                           const int *tab_n, const int *tab_l, int nsep, const double *sep, int nmu, double *d_A,
                           void *stream);

/*
 * Host inputs of coffe_mu_weights_ex.  Background tables on bg_z (as coffe_mu_weights;
 * bg_a = the scale factor, read by the d2 / g terms only), the galaxy biases 1 and 2 on
 * bias_z, the magnification biases s1, s2 and evolution biases f_evo,1/2 on mb_z (d2 / g
 * terms only; Akima like every table, parser.c:807), H0h = COFFE_H0 h in 1/Mpc and
 * Omega_m = Omega0_cdm + Omega0_baryon (g terms), and the renormalisation grid of I_0^4
 * (integrals.c:848-907): nren chi nodes ren_chi and ren[b nren + a] at (ren_chi[a],
 * ren_chi[b]), interpolated by GSL's bicubic (natural-spline node derivatives, the
 * reference's default covariance_interpolation_method, parser.c:810).
 */
typedef struct coffe_mu_inputs {
    int nbg;
    const double *bg_z, *bg_chi, *bg_D1, *bg_f, *bg_cH, *bg_G1, *bg_G2, *bg_a;
    int nb;
    const double *bias_z, *bias1, *bias2;
    int nmb;
    const double *mb_z, *s1, *s2, *fevo1, *fevo2;
    double H0h, Omega_m;
    int nren;
    const double *ren_chi, *ren;
} coffe_mu_inputs;

/*
 * coffe_mu_weights_terms with the d2 and g1-g3 terms and a paired mode (functions.c:
 * 89-1484; no reference counterpart for the pairing).  ntab <= 9 (the 8 standard tables
 * and I_0^4, each at most once).  terms: any COFFE_TERM_* mask;
 * sources: the COFFE_SOURCE_* of the d2 / g terms (0 without them).  Table (n, l) = (4,
 * 0) is the divergent I_0^4 (its r^4 I_0^4 table, integrals.c:772-834) and is needed by
 * COFFE_TERM_DG_DG, which also integrates minus the renormalisation grid at (chi1,
 * chi2) against the same mu weights into d_ren [nz][nl][nsep] (device, required then):
 *   multipoles = sum_t A[z][l][t][r] I_t(sep_r) + ren[z][l][r].
 * paired = 1: point i is (zmean[i], sep[i]), nz == nsep, d_A [nz][nl][ntab] and d_ren
 * [nz][nl] (the redshift-averaged multipoles' node pairs).  Asynchronous like
 * coffe_mu_weights; COFFE_EDOMAIN when an end point leaves the chi table or the grid.
 */
int coffe_mu_weights_ex(coffe_engine *e, const coffe_mu_inputs *in, unsigned terms, unsigned sources, int flatsky,
                        int paired, int nz, const double *zmean, int nl, const int *ls, int ntab, const int *tab_n,
                        const int *tab_l, int nsep, const double *sep, int nmu, double *d_A, double *d_ren,
                        void *stream);

/*
 * The weights of coffe_mu_weights_ex (density, RSD and Doppler-1 terms, full or flat sky,
 * GL nmu) and their tangents d_At (layout of d_A) along one direction of the background
 * (no reference counterpart: the derivative of the multipoles with respect to the
 * cosmological parameters through the background, functions.c:98-127 lookups, for
 * d M / d theta = sum_t At I_t): tangent->bg_chi, bg_D1, bg_f, bg_cH, bg_G1, bg_G2 are the
 * derivatives of those tables at the nodes in->bg_z (the z grid is fixed; chi's tangent
 * also moves the nodes of the z(chi) inversion), from coffe_background_jvp; the other
 * fields of tangent are ignored and the biases have no tangent.  Forward-mode dual
 * numbers through the Akima lookups (their slopes' |.| included) and the full-sky
 * geometry.  The d2 / g terms return COFFE_EUNSUPPORTED.  Asynchronous like
 * coffe_mu_weights_ex.
 */
int coffe_mu_weights_jvp(coffe_engine *e, const coffe_mu_inputs *in, const coffe_mu_inputs *tangent, unsigned terms,
                         int flatsky, int paired, int nz, const double *zmean, int nl, const int *ls, int ntab,
                         const int *tab_n, const int *tab_l, int nsep, const double *sep, int nmu, double *d_A,
                         double *d_At, void *stream);

/*
 * coffe_mu_weights_ex with an explicit mu rule of nmu >= 1 nodes mu[q] in [-1, 1] and
 * weights wmu[q] in place of the Gauss-Legendre one (A = (2l+1)/2 sum_q wmu_q P_l(mu_q)
 * G c_t): one node mu with weight 2 and l = 0 gives the local correlation function at
 * mu (functions_nonintegrated, functions.c:89-1484, the reference's CORRFUNC output).
 */
int coffe_mu_weights_rule(coffe_engine *e, const coffe_mu_inputs *in, unsigned terms, unsigned sources, int flatsky,
                          int paired, int nz, const double *zmean, int nl, const int *ls, int ntab, const int *tab_n,
                          const int *tab_l, int nsep, const double *sep, int nmu, const double *mu,
                          const double *wmu, double *d_A, double *d_ren, void *stream);

/*
 * The reference's full table (coffe_integrals_renormalizable, integrals.c:459-573):
 * out_x/out_y (caller-allocated, N + 43 entries) = [r = 0, the 42 min_sep values
 * (integrals.c:111-126), the N FFTLog nodes], y = I_l^n(r), times r^(n-l) when n >= l.
 * The small-r nodes are QAG-GK61 integrals (epsrel 1e-5) over [kmin, kmax] of the
 * Akima spline of (k, pk); the FFT nodes are twofast_1bessel(l, nu = n, r0 = 1,
 * k0 = kmin).  Integer n in [-2, 4], l in [0, 4].  Synchronous, default engine.
 */
/*
 * The divergent table I_0^4 and its renormalisations (integrals.c:772-909, the
 * reference's `divergent` branch, built when d2 or g1-g5 is requested), each a
 * QAG-GK61 integral (epsrel 1e-5, integrals.c:331-374) over [kmin, kmax] of the Akima
 * spline P of (k, pk), divided by 2 pi^2:
 *   out_i40[i] = sep_i^4 int dk k^2 P j_0(k sep_i)/(k sep_i)^4   (integrals.c:797-811)
 *   out_r0[i]  = sep_i^4 int dk P (1 - j_0(k sep_i)^2)/k^2       (integrals.c:813-825,
 *                the same sep^(n-l) factor of integrals.c:351-356; the reference then
 *                replaces the first node by (0, 0) for its spline, 838-839)
 *   out_ren[a nchi + b] = int dk P j_0(k chi_b) j_0(k chi_a)/k^2 (integrals.c:883-895,
 *                the (nbins+1)^2 grid result2d[k (nbins+1) + j])
 * sep >= 0 (nsep), chi >= 0 (nchi <= 4096), host arrays; one wavefront per integral
 * (intervals capped at 1024 per integral).  out_status (or NULL), 2 nsep + nchi^2
 * entries in the order i40, r0, ren: QUADPACK's code (0 converged, 1 interval cap,
 * 2 roundoff, 3 bad integrand behaviour) or 4 for a non-finite value; out_intervals
 * (or NULL), same layout: the final number of intervals (GSL's workspace size; the
 * rule was evaluated 61 (2 n - 1) times).  As the
 * reference (its GSL error handler is off here, integrals.c:593) roundoff-limited
 * results are kept; returns COFFE_ENONFINITE for a non-finite value and COFFE_ENOCONV
 * when an integral reached the interval cap, the values still written.  Synchronous,
 * default engine.
 */
int coffe_divergent_tables(const double *k, const double *pk, size_t nin, double kmin, double kmax,
                           const double *sep, size_t nsep, const double *chi, size_t nchi, double *out_i40,
                           double *out_r0, double *out_ren, int *out_status, int *out_intervals);
/* the same on device arrays, asynchronous on stream (NULL = the engine's): d_status
   (required) and d_intervals (or NULL) as above, the caller reads them after the
   stream; [kmin, kmax] must lie inside [d_k[0], d_k[nin-1]] (outside, the Akima
   spline is NaN and the integral reports status 4) */
int coffe_divergent_tables_device(coffe_engine *e, const double *d_k, const double *d_pk, size_t nin, double kmin,
                                  double kmax, const double *d_sep, size_t nsep, const double *d_chi, size_t nchi,
                                  double *d_i40, double *d_r0, double *d_ren, int *d_status, int *d_intervals,
                                  void *stream);

int coffe_integral_table(int n, int l, size_t N, const double *k, const double *pk, size_t nin, double kmin,
                         double kmax, double *out_x, double *out_y);

/*
 * The same table with the reference's integer states (integrals.c:186-207, enum
 * coffe_integer_state of integrals.h:23: 0 COFFE_INTEGER, 1 COFFE_HALF_INTEGER): with
 * both states 1 the table is that of (n/2, l/2) for odd n in [-3, 9], l in [-1, 9], the
 * half-integer tables of the flat-sky lensing terms (integrals.c:917-960, 1246-1300):
 * FFT nodes from twofast_1bessel with half-integer l and nu, small-r nodes from the
 * kernel sqrt(pi/2) J_{l/2+1/2}(kr) / (kr)^{n/2+1/2} (integrals.c:296-306) and
 * I(0) = sqrt(pi/2) int k^{2+l/2-n/2} P / (2^m m!) / (2 pi^2), m = l/2 + 1/2
 * (integrals.c:215-231).  Mixed states return COFFE_EUNSUPPORTED.
 */
int coffe_integral_table_state(int n, int l, int state_n, int state_l, size_t N, const double *k, const double *pk,
                               size_t nin, double kmin, double kmax, double *out_x, double *out_y);

/*
 * The 43 small-separation nodes (r = 0 and the 42 min_sep values) of every table of
 * a batch, for its spectra: d_out[nspec][ntab][43] (device).  Synchronises the
 * stream and returns COFFE_ENOCONV when an integral did not converge.
 */
int coffe_batch_small_r(coffe_batch *b, const double *d_k, const double *d_pk, double *d_out, void *stream);

/*
 * Forward complex FFT of any length L (no reference counterpart: the reference calls
 * FFTW): out_k = sum_n in_n e^{-2 pi i k n / L}, unnormalised, interleaved (re, im)
 * host arrays of 2L doubles.  The device transform behind the FFTLog drop-in for
 * lengths outside the fused LDS kernels: one LDS transform for 2^a 3^b 5^c <= 8192,
 * four-step beyond, Bluestein for other factors.  L <= 2^26; synchronous, default
 * engine.
 */
int coffe_fft_any(const double *in, double *out, int64_t L);

/*
 * Flat-sky lensing multipoles, the two lensing terms the reference computes without
 * Monte Carlo when its flat-sky switches are on (signal.c:612-628, 704-736):
 * lensing-lensing (functions.c:4020-4133, flatsky_nonlocal; the x integral over
 * [0, 1] by a composite Gauss-Legendre rule of `panels` panels of 32 nodes in place
 * of the reference's QAG) and density-lensing (functions.c:4136-4257,
 * flatsky_local_nonlocal), each already multiplied by 2 pi^2 as signal.c adds them,
 * summed into out[nz][nl][nsep] when want_ll / want_dl is set.  Odd multipoles give 0.
 *   zpar[nz][7]   z_mean, chi(z_mean), D1(z_mean), s1, s2 (magnification biases) and
 *                 b1, b2 (galaxy biases) at z_mean;
 *   ls[nl], sep[nsep] (Mpc); bg_z, bg_chi, bg_D1[nb] the background samples (chi
 *                 increasing: the reference's z_as_chi spline is Akima of z against chi);
 *   tab_x, tab_y  the tables, concatenated: the (n, l) = (1, l) table of every even
 *                 multipole at ll_off[il] (length ll_len[il], 0 for none) and the
 *                 half-integer tables ((l - 2k + 3)/2, (l - 2k + 1)/2) at
 *                 dl_off[il * 5 + k] (lengths dl_len, k <= l/2), as
 *                 coffe_integral_table / twofast_1bessel give them (r^(n-l) I_l^n for
 *                 n >= l), at least 5 nodes each;
 *   Omega_m, H0h  Omega_cdm + Omega_baryon and H0 h in 1/Mpc.
 * Multipoles up to 8 with want_dl.  Synchronous on the default engine; COFFE_EINVAL
 * for bad sizes, NaN where a lookup falls outside its table.
 */
int coffe_flat_lensing_multipoles(int nz, const double *zpar, int nl, const int *ls, int nsep, const double *sep,
                                  int nb, const double *bg_z, const double *bg_chi, const double *bg_D1,
                                  int64_t ntab_values, const double *tab_x, const double *tab_y, const int64_t *ll_off,
                                  const int *ll_len, const int64_t *dl_off, const int *dl_len, double Omega_m,
                                  double H0h, int want_ll, int want_dl, int panels, double *out);

/*
 * The flat-sky lensing correlation function at every (z_mean, mu, separation), the
 * branches the reference takes outside MULTIPOLES with its flat-sky switches
 * (functions.c:3515-3558, lensing-lensing with flatsky_nonlocal, whose second
 * line-of-sight variable does not enter; functions.c:1563-1622, density-lensing with
 * flatsky_local_nonlocal): with T(r) the (n, l) = (1/2, -1/2) table times
 * sqrt(2/pi) 2 pi^2 (integrals.c:956-962) and s = sqrt(1 - mu^2),
 *   lensing-lensing = int_0^1 dx 9 Omega_m^2 (2 - 5 s1)(2 - 5 s2) chi^3 / (8 pi) T(x r s)
 *                     D1(z(x chi))^2 (1 + z(x chi))^2 (x (1 - x))^2 (H0 h)^4,
 *   density-lensing = (H0 h)^2 (-3 Omega_m / (8 pi)) D1^2 (1 + z) |mu| r
 *                     ((2 - 5 s1) b2 + (2 - 5 s2) b1) T(r s),
 * summed into out[nz][nmu][nsep] when want_ll / want_dl is set.  tab_x, tab_y is the
 * table as coffe_integral_table_state(1, -1, 1, 1) gives it (ntab >= 5 nodes, r = 0
 * first); zpar, bg_*, Omega_m, H0h and panels as for coffe_flat_lensing_multipoles;
 * mu in [-1, 1].  Synchronous on the default engine.
 */
int coffe_flat_lensing_corrfunc(int nz, const double *zpar, int nmu, const double *mu, int nsep, const double *sep,
                                int nb, const double *bg_z, const double *bg_chi, const double *bg_D1, int ntab,
                                const double *tab_x, const double *tab_y, double Omega_m, double H0h, int want_ll,
                                int want_dl, int panels, double *out);

/* contributions of coffe_integrated_corrfunc (the reference's correlation_contrib
   switches, common.h, and its flat-sky switches) */
enum {
    COFFE_INTEGRATED_DEN = 1u << 0,
    COFFE_INTEGRATED_RSD = 1u << 1,
    COFFE_INTEGRATED_D1 = 1u << 2,
    COFFE_INTEGRATED_D2 = 1u << 3,
    COFFE_INTEGRATED_G1 = 1u << 4,
    COFFE_INTEGRATED_G2 = 1u << 5,
    COFFE_INTEGRATED_G3 = 1u << 6,
    COFFE_INTEGRATED_LEN = 1u << 7,
    COFFE_INTEGRATED_G4 = 1u << 8,
    COFFE_INTEGRATED_G5 = 1u << 9,
    COFFE_INTEGRATED_ONLY_CROSS = 1u << 10,   /* only_cross_correlations: no auto terms */
    COFFE_INTEGRATED_FLAT_LL = 1u << 11,      /* flatsky_nonlocal: lensing-lensing left to the flat-sky forms */
    COFFE_INTEGRATED_FLAT_DL = 1u << 12       /* flatsky_local_nonlocal: density/RSD/d1-lensing likewise */
};

/*
 * The integrated terms of the correlation function, full sky, at every (z_mean, mu,
 * separation): the single-integrated terms (functions_single_integrated, functions.c:
 * 1486-3458: density, RSD, d1, d2, g1-g3 against lensing, g4, g5) and the double-
 * integrated ones (functions_double_integrated, functions.c:3460-4017: lensing, g4 and
 * g5 with each other), integrated over the line-of-sight variables as coffe_integrate
 * does for CORRFUNC (signal.c:556-741), with deterministic product rules in place of
 * the reference's QAG / double-exponential and CUBA / Monte Carlo integrators.
 * out[((iz nmu + im) nsep + ir) 2 + {0, 1}] = (single, double).
 *   in        background (bg_a required), galaxy biases, magnification and evolution
 *             biases, Omega_m, H0 h, and with g4 / g5 the renormalisation grid of I_0^4
 *             (integrals.c:848-907) as for coffe_mu_weights_ex;
 *   flags     COFFE_INTEGRATED_* (the contributions present);
 *   tables    ntab tables (tab_n, tab_l), concatenated in tab_x / tab_y with lengths
 *             tab_len: the (n, l) of (0,0), (0,2), (0,4), (1,1), (1,3), (2,0), (2,2),
 *             (3,1) as coffe_integral_table gives them and (4, 0) the divergent r^4
 *             I_0^4 table; those the flags need must be present (5 nodes or more);
 *   ren0      with g4 / g5: the r = 0 renormalisation of I_0^4 against separation
 *             (integrals.c:829-835, first node (0, 0));
 *   rule      {p1, n1, pu, nu, ps, ns} or NULL for {8, 16, 12, 16, 32, 16}: x1 as p1
 *             panels of n1 Gauss-Legendre nodes, the sinh-mapped x2 as pu panels of
 *             nu, each single-integrated piece as ps panels of ns.
 * Synchronous on the default engine; COFFE_EINVAL for bad sizes, flags or a missing
 * table, COFFE_EDOMAIN for |mu| > 1 - 1e-10 (along the line of sight the separations
 * of the integration points fall to FP64 rounding noise and into the terms' r = 0
 * branches), NaN where a lookup leaves its table.
 */
int coffe_integrated_corrfunc(const coffe_mu_inputs *in, unsigned flags, int ntab, const int *tab_n, const int *tab_l,
                              const int *tab_len, const double *tab_x, const double *tab_y, int nren0,
                              const double *ren0_x, const double *ren0_y, int nz, const double *zmean, int nmu,
                              const double *mu, int nsep, const double *sep, const int *rule, double *out);

/* synchronise the engine's stream */
int coffe_engine_synchronize(coffe_engine *e);

#ifdef __cplusplus
}
#endif
#endif
