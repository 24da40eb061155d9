/*
 * Oracle (test infrastructure only): CPU restatement of COFFE's 1D FFTLog and the
 * I_l^n(r) table builder, following the reference step by step.
 *
 *   orc_twofast_window      <- twofast.c:54-90
 *   orc_twofast_lgamma      <- twofast.c:97-144   (Lanczos g = 7, 8 coefficients,
 *                                                   reflection for Re z < 0.5)
 *   orc_twofast_mql         <- twofast.c:196-224
 *   orc_twofast_selectqnu   <- twofast.c:240-250
 *   fft_input (static)      <- twofast.c:252-360  (C branch, 345-353)
 *   orc_twofast_1bessel     <- twofast.c:362-483
 *   orc_integral_table      <- integrals.c:459-573 (+ min_sep 111-126,
 *                                                   coefficients 215-231,
 *                                                   integrands 238-325, 333-383)
 * cpow(a, w) is evaluated as cexp(w clog(a)), the C99 library definition.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef double complex cplx;

#define ORC_PI 3.14159265358979323846
#define ORC_SQRT_PI 1.7724538509055160272981674833411451827975494561224

int orc_twofast_window(double value, double xmin, double xmax, double xleft, double xright, double *out)
{
    if (!(xmin <= xleft && xleft <= xright && xright <= xmax)) return ORC_EDOMAIN;
    double result;
    if (value > xleft && value < xright && value > xmin && value < xmax) {
        result = 1.0;
    } else if (value <= xmin || value >= xmax) {
        result = 0.0;
    } else {
        /* the reference leaves value == xleft / xright undefined; the ramp's
           continuous value (1 - sin(2 pi)/2 pi) is used there */
        if (value <= xleft) result = (value - xmin) / (xleft - xmin);
        else result = (xmax - value) / (xmax - xright);
        result = result - sin(2 * ORC_PI * result) / 2. / ORC_PI;
    }
    *out = result;
    return ORC_OK;
}

static const double LANCZOS7[8] = {
    676.5203681218851, -1259.1392167224028, 771.32342877765313, -176.61502916214059,
    12.507343278686905, -0.13857109526572012, 9.9843695780195716e-6, 1.5056327351493116e-7};

static cplx lgamma_c(cplx z)
{
    if (creal(z) < 0.5) return log(ORC_PI) - clog(csin(ORC_PI * z)) - lgamma_c(1. - z);
    cplx x = 0.99999999999980993;
    z -= 1;
    for (int i = 0; i < 8; ++i) x += LANCZOS7[i] / (z + (double)i + 1.);
    cplx t = z + 8.0 - 0.5;
    return log(2 * ORC_PI) / 2. + (z + 0.5) * clog(t) - t + clog(x);
}

void orc_twofast_lgamma(double re, double im, double *out)
{
    cplx r = lgamma_c(re + I * im);
    out[0] = creal(r);
    out[1] = cimag(r);
}

static cplx cpow_real_base(double a, cplx w)
{
    return cexp(w * clog((cplx)a));
}

static int mql(double t, double q, double l, double alpha, cplx *out)
{
    const cplx n = q - 1. - t * I;
    cplx unl = cpow_real_base(2.0, n - 1.) * ORC_SQRT_PI *
               cexp(lgamma_c((1. + l + n) / 2.) - lgamma_c((2. + l - n) / 2.));
    if (!(isfinite(creal(unl)) && isfinite(cimag(unl)))) return ORC_ENONFINITE;
    *out = cpow_real_base(alpha, t * I - q) * unl;
    return ORC_OK;
}

int orc_twofast_mql(double t, double q, double l, double alpha, double *out)
{
    cplx r = 0;
    int st = mql(t, q, l, alpha, &r);
    out[0] = creal(r);
    out[1] = cimag(r);
    return st;
}

double orc_twofast_selectqnu(double l, double nu)
{
    const double n1 = 0.9, n2 = 0.9999;
    double qmin = fmax(n2 - 1.0 - nu, -l);
    double qmax = fmin(n1 + 3.0 - nu, 2.0);
    double q = (2. + n1 + n2) / 2. - nu;
    if (!(qmin < q && q < qmax)) q = (qmin + 2. * qmax) / 3.;
    return q;
}

/* twofast.c:252-360; out: N/2+1 complex (interleaved) */
static int fft_input(cplx *output_y, size_t N, size_t nin, const double *kin, const double *pin,
                     const double *c, double q, double k0, double kmin, double kmax)
{
    const size_t N2 = N / 2 + 1;
    const double L = 2 * ORC_PI * N / log(kmax / kmin);
    double *xm = (double *)malloc(sizeof(double) * N);
    double *ym = (double *)malloc(sizeof(double) * N);
    double *Y = (double *)malloc(sizeof(double) * 2 * N2);
    orc_rfft *plan = orc_rfft_new(N);
    if (!xm || !ym || !Y || !plan) { free(xm); free(ym); free(Y); orc_rfft_free(plan); return ORC_ENOMEM; }
    const double xmax = k0 * pow(kmax / kmin, (double)(N - 1) / N);
    const double xl = exp(0.46) * kmin, xr = exp(-0.46) * xmax;
    int st = ORC_OK;
    for (size_t i = 0; i < N; ++i) xm[i] = k0 * pow(kmax / kmin, (double)i / N);
    for (size_t i = 0; i < N && st == ORC_OK; ++i) {
        const double x = k0 * pow(kmax / kmin, (double)i / N);
        double w;
        st = orc_twofast_window(x, kmin, xmax, xl, xr, &w);
        ym[i] = pow(kmax / kmin, (3. - q) * i / N) * orc_cspline_eval(nin, kin, pin, c, x) * w;
    }
    if (st == ORC_OK) {
        orc_rfft_forward(plan, ym, Y);
        for (size_t i = 0; i < N2 && st == ORC_OK; ++i) {
            double w;
            st = orc_twofast_window(xm[N2 - 2 + i], kmin, xmax, xl, xr, &w);
            output_y[i] = w * conj(Y[2 * i] + I * Y[2 * i + 1]) / L;
        }
    }
    free(xm); free(ym); free(Y); orc_rfft_free(plan);
    return st;
}

int orc_twofast_1bessel(double *output_x, double *output_y, size_t N, const double *input_x,
                        const double *input_y, size_t input_len, double l, double nu, double r0, double k0,
                        double kmin, double kmax)
{
    if (N < 2 || (N & 1) || input_len < 3 || !(kmax > kmin) || !(kmin > 0)) return ORC_EINVAL;
    for (size_t i = 0; i < input_len; ++i)
        if (!isfinite(input_x[i]) || !isfinite(input_y[i])) return ORC_ENONFINITE;
    const size_t N2 = N / 2 + 1;
    const double qnu = orc_twofast_selectqnu(l, nu);
    const double G = log(kmax / kmin);
    double *c = (double *)malloc(sizeof(double) * input_len);
    cplx *Y = (cplx *)malloc(sizeof(cplx) * N2);
    double *X = (double *)malloc(sizeof(double) * 2 * N2);
    double *pref = (double *)malloc(sizeof(double) * N);
    double *tmp = (double *)malloc(sizeof(double) * N);
    orc_rfft *plan = orc_rfft_new(N);
    int st = ORC_ENOMEM;
    if (!c || !Y || !X || !pref || !tmp || !plan) goto done;
    st = orc_cspline_init(input_len, input_x, input_y, c);
    if (st != ORC_OK) goto done;
    st = fft_input(Y, N, input_len, input_x, input_y, c, qnu + nu, k0, kmin, kmax);
    if (st != ORC_OK) goto done;
    for (size_t i = 0; i < N; ++i) output_x[i] = r0 * pow(kmax / kmin, (double)i / N);
    for (size_t i = 0; i < N; ++i)
        pref[i] = k0 * k0 * k0 * pow(kmax / kmin, -((qnu + nu) * i / N)) / ORC_PI / pow(r0 * k0, nu) / G;
    for (size_t i = 0; i < N2; ++i) {
        cplx u;
        st = mql(2 * ORC_PI * i / G, qnu, l, k0 * r0, &u);
        if (st != ORC_OK) goto done;
        cplx v = Y[i] * u;
        X[2 * i] = creal(v);
        X[2 * i + 1] = cimag(v);
    }
    orc_rfft_backward(plan, X, tmp);
    for (size_t i = 0; i < N; ++i) output_y[i] = tmp[i] * pref[i];
    st = ORC_OK;
done:
    free(c); free(Y); free(X); free(pref); free(tmp); orc_rfft_free(plan);
    return st;
}

/* ---- I_l^n(r) tables (integrals.c) --------------------------------------- */

const double ORC_MIN_SEP[42] = {
    1E-6, 1E-5, 1E-4, 1E-3, 2E-3, 5E-3, 7E-3, 8E-3, 9E-3, 1E-2, 1.1E-2, 1.2E-2, 1.25E-2, 1.3E-2,
    1.35E-2, 1.5E-2, 2E-2, 2.5E-2, 3E-2, 3.5E-2, 5E-2, 7E-2, 8E-2, 9E-2, 1E-1, 1.1E-1, 1.2E-1,
    1.25E-1, 1.3E-1, 1.35E-1, 1.5E-1, 1.75E-1, 2E-1, 2.5E-1, 3E-1, 3.5E-1, 4E-1, 5E-1, 6E-1,
    7E-1, 8E-1, 9E-1};

typedef struct {
    size_t nin;
    const double *k, *p, *b, *c, *d;   /* Akima spline of P(k) */
    double r, n, l;
} bessel_params;

/* k^(2 + l - n) P(k)  (integrals.c:238-269, integer states) */
static double integrand_prefactor(double k, void *vp)
{
    const bessel_params *p = (const bessel_params *)vp;
    return pow(k, 2 + p->l - p->n) * orc_akima_eval(p->nin, p->k, p->p, p->b, p->c, p->d, k);
}

/* k^2 P(k) j_l(k r)/(k r)^n  (integrals.c:280-325, integer states) */
static double integrand_bessel(double k, void *vp)
{
    const bessel_params *p = (const bessel_params *)vp;
    const double bes = orc_sph_jl((int)p->l, k * p->r) / pow(k * p->r, p->n);
    return k * k * orc_akima_eval(p->nin, p->k, p->p, p->b, p->c, p->d, k) * bes;
}

/* sqrt(pi/2) / 2^(l+1/2) / Gamma(l + 3/2) = 1/(2l+1)!!  (integrals.c:215-231) */
static double coefficient_r0(int l)
{
    return sqrt(ORC_PI / 2.) / pow(2, l + 0.5) / tgamma(l + 1.5);
}

int orc_integral_table(int n, int l, size_t N, size_t nin, const double *kin, const double *pin, double kmin,
                       double kmax, int with_small_r, double *out_x, double *out_y)
{
    const size_t nmin = 42;
    double *fx = (double *)malloc(sizeof(double) * N);
    double *fy = (double *)malloc(sizeof(double) * N);
    double *akb = (double *)malloc(sizeof(double) * 3 * nin);
    int st = ORC_ENOMEM;
    if (!fx || !fy || !akb) goto done;
    st = orc_twofast_1bessel(fx, fy, N, kin, pin, nin, (double)l, (double)n, 1, kmin, kmin, kmax);
    if (st != ORC_OK) goto done;
    st = orc_akima_init(nin, kin, pin, akb, akb + nin, akb + 2 * nin);
    if (st != ORC_OK) goto done;
    bessel_params bp = {nin, kin, pin, akb, akb + nin, akb + 2 * nin, 0.0, (double)n, (double)l};
    double x0 = 0.0;
    if (n >= l) {
        for (size_t i = 0; i < N; ++i) fy[i] = fy[i] * pow(fx[i], (double)(n - l));
        if (with_small_r) {
            const double res = orc_qag(integrand_prefactor, &bp, kmin, kmax, 0, 1e-5, 50000, NULL, NULL);
            x0 = res / 2. / ORC_PI / ORC_PI * coefficient_r0(l);
        }
    }
    out_x[0] = 0.0;
    out_y[0] = x0;
    for (size_t i = 1; i <= nmin; ++i) {
        out_x[i] = ORC_MIN_SEP[i - 1];
        out_y[i] = 0.0;
        if (with_small_r) {
            bp.r = out_x[i];
            double res = orc_qag(integrand_bessel, &bp, kmin, kmax, 0, 1e-5, 50000, NULL, NULL);
            if (n >= l) res *= pow(out_x[i], (double)(n - l));
            out_y[i] = res / 2. / ORC_PI / ORC_PI;
        }
    }
    for (size_t i = 0; i < N; ++i) {
        out_x[nmin + 1 + i] = fx[i];
        out_y[nmin + 1 + i] = fy[i];
    }
    st = ORC_OK;
done:
    free(fx); free(fy); free(akb);
    return st;
}

/* the divergent table I_0^4 and its renormalisations (integrals.c:772-909): QAG-GK61
   at epsrel 1e-5 with a 50000-interval workspace (common.c:1004-1040) of
     k^2 P j_0(k r)/(k r)^4                  times r^4      (integrals.c:280-325, 351-356)
     P (1 - j_0(k r)^2)/k^2                  times r^4      (integrals.c:389-398, 351-356)
     P j_0(k chi_1) j_0(k chi_2)/k^2                        (integrals.c:405-448)
   each divided by 2 pi^2; ren[a nchi + b] = (chi_b, chi_a) as result2d[k (nbins+1) + j]
   (integrals.c:883-895).  OpenMP over the integrals. */
typedef struct {
    bessel_params bp;
    double r2;
} renorm_params;

static double integrand_renorm0(double k, void *vp)
{
    const bessel_params *p = (const bessel_params *)vp;
    const double j0 = orc_sph_jl(0, k * p->r);
    return orc_akima_eval(p->nin, p->k, p->p, p->b, p->c, p->d, k) * (1. - j0 * j0) / k / k;
}

static double integrand_renorm(double k, void *vp)
{
    const renorm_params *q = (const renorm_params *)vp;
    const bessel_params *p = &q->bp;
    const double jj = orc_sph_jl(0, k * p->r) * orc_sph_jl(0, k * q->r2);
    return orc_akima_eval(p->nin, p->k, p->p, p->b, p->c, p->d, k) * jj / k / k;
}

int orc_divergent_tables(size_t nin, const double *kin, const double *pin, double kmin, double kmax, size_t nsep,
                         const double *sep, size_t nchi, const double *chi, double *i40, double *r0, double *ren)
{
    double *akb = (double *)malloc(sizeof(double) * 3 * nin);
    if (!akb) return ORC_ENOMEM;
    int st = orc_akima_init(nin, kin, pin, akb, akb + nin, akb + 2 * nin);
    if (st != ORC_OK) { free(akb); return st; }
    const long long units = (long long)(2 * nsep + nchi * nchi);
    #pragma omp parallel for schedule(dynamic, 1)
    for (long long u = 0; u < units; ++u) {
        renorm_params q;
        bessel_params bp = {nin, kin, pin, akb, akb + nin, akb + 2 * nin, 0.0, 4.0, 0.0};
        q.bp = bp;
        q.r2 = 0.0;
        double v;
        if (u < (long long)(2 * nsep)) {
            const int direct = u < (long long)nsep;
            const double r = sep[direct ? u : u - (long long)nsep];
            q.bp.r = r;
            v = orc_qag(direct ? integrand_bessel : integrand_renorm0, &q.bp, kmin, kmax, 0, 1e-5, 50000, NULL, NULL);
            if (r > 0) v *= pow(r, 4.0);
            v = v / 2. / ORC_PI / ORC_PI;
            if (direct) i40[u] = v; else r0[u - (long long)nsep] = v;
        } else {
            const long long w = u - (long long)(2 * nsep);
            q.bp.r = chi[w % (long long)nchi];
            q.r2 = chi[w / (long long)nchi];
            v = orc_qag(integrand_renorm, &q, kmin, kmax, 0, 1e-5, 50000, NULL, NULL);
            ren[w] = v / 2. / ORC_PI / ORC_PI;
        }
    }
    free(akb);
    return ORC_OK;
}

/* tables at separations for a batch: out[s][t][r] = Akima(table_t of spectrum s)(sep_r),
   tables without the small-r nodes (they do not enter Akima stencils at sep >= r_2);
   OpenMP over (spectrum, table) */
int orc_tables_at_sep_batch(int nspec, const long long *off, const double *k, const double *p, const long long *nfft,
                            int ntab, const int *tab_n, const int *tab_l, size_t nsep, const double *sep, double *out)
{
    int status = ORC_OK;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int u = 0; u < nspec * ntab; ++u) {
        const int s = u / ntab, t = u % ntab;
        const size_t nin = (size_t)(off[s + 1] - off[s]);
        const size_t N = (size_t)nfft[s];
        double *tx = (double *)malloc(sizeof(double) * (N + 43));
        double *ty = (double *)malloc(sizeof(double) * (N + 43));
        int st = (tx && ty) ? ORC_OK : ORC_ENOMEM;
        if (st == ORC_OK)
            st = orc_integral_table(tab_n[t], tab_l[t], N, nin, k + off[s], p + off[s], k[off[s]], k[off[s + 1] - 1],
                                    0, tx, ty);
        if (st == ORC_OK) st = orc_akima_interp(N + 43, tx, ty, nsep, sep, out + (size_t)u * nsep);
        free(tx);
        free(ty);
        if (st != ORC_OK) {
            #pragma omp atomic write
            status = st;
        }
    }
    return status;
}

/*
 * Forward-mode derivative of orc_tables_at_sep_batch along a direction v of the input
 * samples (same layout as p): the tables (FFTLog of the spline of the samples,
 * twofast.c:362-483) are linear in the samples, so their tangent is the table of v;
 * the Akima lookups at the separations (common.c:704-710) are differentiated by
 * orc_akima_jvp.  out as orc_tables_at_sep_batch, dout the tangent, both
 * [nspec][ntab][nsep].
 */
int orc_tables_at_sep_jvp(int nspec, const long long *off, const double *k, const double *p, const double *v,
                          const long long *nfft, int ntab, const int *tab_n, const int *tab_l, size_t nsep,
                          const double *sep, double *out, double *dout)
{
    int status = ORC_OK;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int u = 0; u < nspec * ntab; ++u) {
        const int s = u / ntab, t = u % ntab;
        const size_t nin = (size_t)(off[s + 1] - off[s]);
        const size_t N = (size_t)nfft[s];
        double *tx = (double *)malloc(sizeof(double) * (N + 43));
        double *ty = (double *)malloc(sizeof(double) * (N + 43));
        double *vx = (double *)malloc(sizeof(double) * (N + 43));
        double *vy = (double *)malloc(sizeof(double) * (N + 43));
        int st = (tx && ty && vx && vy) ? ORC_OK : ORC_ENOMEM;
        const double kmin = k[off[s]], kmax = k[off[s + 1] - 1];
        if (st == ORC_OK) st = orc_integral_table(tab_n[t], tab_l[t], N, nin, k + off[s], p + off[s], kmin, kmax, 0, tx, ty);
        if (st == ORC_OK) st = orc_integral_table(tab_n[t], tab_l[t], N, nin, k + off[s], v + off[s], kmin, kmax, 0, vx, vy);
        if (st == ORC_OK)
            st = orc_akima_jvp(N + 43, tx, ty, vy, nsep, sep, out + (size_t)u * nsep, dout + (size_t)u * nsep);
        free(tx); free(ty); free(vx); free(vy);
        if (st != ORC_OK) {
            #pragma omp atomic write
            status = st;
        }
    }
    return status;
}
