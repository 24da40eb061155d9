/*
 * Oracle (test infrastructure only): quadrature rules and special functions.
 *
 *  - QUADPACK QAG with the 61-point Gauss-Kronrod rule, as COFFE calls it through
 *    GSL (common.c:1004-1040: epsabs 0, epsrel = `integration_1d_prec`, limit =
 *    COFFE_MAX_INTSPACE = 50000, GSL_INTEG_GAUSS61).  Error estimate rescaled as
 *    QUADPACK does: err = resasc * min(1, (200 |K-G| / resasc)^1.5), floored at
 *    50 eps resabs; the interval with the largest error is bisected until the
 *    summed error meets max(epsabs, epsrel |area|) or roundoff is detected.
 *  - Gauss-Legendre nodes/weights by Newton iteration on the three-term
 *    recurrence (used for the mu-quadrature pin and as an exact-rule cross check).
 *  - spherical Bessel j_l(x) for integer l >= 0 (GSL sf_bessel_jl is called at
 *    integrals.c:309-316): power series for small x, upward recurrence from
 *    j_0, j_1 when x > l, otherwise Miller's downward recurrence normalised with
 *    sum_k (2k+1) j_k^2 = 1.
 *  - Legendre polynomials P_l by the Bonnet recurrence (gsl_sf_legendre_Pl at
 *    signal.c:109).
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>

#include "oracle.h"
#include "gk61_table.h"

void orc_qk61(orc_fn f, void *p, double a, double b, double *result, double *abserr, double *resabs,
              double *resasc)
{
    const int n = 31;
    double fv1[31], fv2[31];
    const double center = 0.5 * (a + b);
    const double half = 0.5 * (b - a);
    const double abs_half = fabs(half);
    const double fc = f(center, p);
    double rg = 0.0;
    double rk = fc * GK61_WGK[n - 1];
    double rabs = fabs(rk);
    for (int j = 0; j < (n - 1) / 2; ++j) {
        const int jtw = 2 * j + 1;
        const double ab = half * GK61_XGK[jtw];
        const double f1 = f(center - ab, p), f2 = f(center + ab, p);
        fv1[jtw] = f1; fv2[jtw] = f2;
        rg += GK61_WG[j] * (f1 + f2);
        rk += GK61_WGK[jtw] * (f1 + f2);
        rabs += GK61_WGK[jtw] * (fabs(f1) + fabs(f2));
    }
    for (int j = 0; j < n / 2; ++j) {
        const int jtwm1 = 2 * j;
        const double ab = half * GK61_XGK[jtwm1];
        const double f1 = f(center - ab, p), f2 = f(center + ab, p);
        fv1[jtwm1] = f1; fv2[jtwm1] = f2;
        rk += GK61_WGK[jtwm1] * (f1 + f2);
        rabs += GK61_WGK[jtwm1] * (fabs(f1) + fabs(f2));
    }
    const double mean = 0.5 * rk;
    double rasc = GK61_WGK[n - 1] * fabs(fc - mean);
    for (int j = 0; j < n - 1; ++j) rasc += GK61_WGK[j] * (fabs(fv1[j] - mean) + fabs(fv2[j] - mean));
    double err = (rk - rg) * half;
    rk *= half;
    rabs *= abs_half;
    rasc *= abs_half;
    err = fabs(err);
    if (rasc != 0.0 && err != 0.0) {
        const double scale = pow(200.0 * err / rasc, 1.5);
        err = scale < 1.0 ? rasc * scale : rasc;
    }
    if (rabs > DBL_MIN / (50.0 * DBL_EPSILON)) {
        const double min_err = 50.0 * DBL_EPSILON * rabs;
        if (min_err > err) err = min_err;
    }
    *result = rk;
    *abserr = err;
    *resabs = rabs;
    *resasc = rasc;
}

/* Adaptive bisection.  Intervals are kept in insertion order; the one to split is
   the one with the largest error (first one on ties), as QUADPACK's ordered list
   yields for distinct errors. */
double orc_qag(orc_fn f, void *p, double a, double b, double epsabs, double epsrel, size_t limit,
               double *abserr_out, int *status)
{
    double r0, e0, rabs0, rasc0;
    int st = 0;
    orc_qk61(f, p, a, b, &r0, &e0, &rabs0, &rasc0);
    double tol = fmax(epsabs, epsrel * fabs(r0));
    const double round_off = 50.0 * DBL_EPSILON * rabs0;
    if (abserr_out) *abserr_out = e0;
    if (e0 <= round_off && e0 > tol) { if (status) *status = 2; return r0; }
    if ((e0 <= tol && e0 != rasc0) || e0 == 0.0) { if (status) *status = 0; return r0; }
    if (limit == 1) { if (status) *status = 1; return r0; }
    double *ia = (double *)malloc(sizeof(double) * limit * 4);
    if (!ia) { if (status) *status = -1; return NAN; }
    double *ib = ia + limit, *ir = ib + limit, *ie = ir + limit;
    size_t size = 1;
    ia[0] = a; ib[0] = b; ir[0] = r0; ie[0] = e0;
    double area = r0, errsum = e0;
    int rt1 = 0, rt2 = 0;
    size_t iter = 1;
    size_t imax = 0;
    do {
        const double a1 = ia[imax], b2 = ib[imax];
        const double b1 = 0.5 * (a1 + b2), a2 = b1;
        const double ri = ir[imax], ei = ie[imax];
        double r1, e1, ra1, rs1, r2, e2, ra2, rs2;
        orc_qk61(f, p, a1, b1, &r1, &e1, &ra1, &rs1);
        orc_qk61(f, p, a2, b2, &r2, &e2, &ra2, &rs2);
        const double area12 = r1 + r2, err12 = e1 + e2;
        errsum += err12 - ei;
        area += area12 - ri;
        if (rs1 != e1 && rs2 != e2) {
            const double delta = ri - area12;
            if (fabs(delta) <= 1.0e-5 * fabs(area12) && err12 >= 0.99 * ei) rt1++;
            if (iter >= 10 && err12 > ei) rt2++;
        }
        tol = fmax(epsabs, epsrel * fabs(area));
        if (errsum > tol) {
            if (rt1 >= 6 || rt2 >= 20) st = 2;
            {
                const double tmp = (1 + 100 * DBL_EPSILON) * (fabs(a2) + 1000 * DBL_MIN);
                if (fabs(a1) <= tmp && fabs(b2) <= tmp) st = 3;
            }
        }
        /* the bisected interval keeps the half with the larger error */
        if (e2 > e1) {
            ia[imax] = a2; ib[imax] = b2; ir[imax] = r2; ie[imax] = e2;
            ia[size] = a1; ib[size] = b1; ir[size] = r1; ie[size] = e1;
        } else {
            ia[imax] = a1; ib[imax] = b1; ir[imax] = r1; ie[imax] = e1;
            ia[size] = a2; ib[size] = b2; ir[size] = r2; ie[size] = e2;
        }
        size++;
        imax = 0;
        for (size_t k = 1; k < size; ++k) if (ie[k] > ie[imax]) imax = k;
        iter++;
    } while (iter < limit && !st && errsum > tol);
    double sum = 0.0;
    for (size_t k = 0; k < size; ++k) sum += ir[k];
    free(ia);
    if (abserr_out) *abserr_out = errsum;
    if (status) *status = (errsum <= tol) ? 0 : (st ? st : 1);
    return sum;
}

int orc_gauss_legendre(size_t n, double *x, double *w)
{
    /* nodes on [-1, 1], ascending; long double Newton for full double accuracy */
    for (size_t i = 0; i < (n + 1) / 2; ++i) {
        long double z = cosl(3.14159265358979323846264338327950288L * ((long double)i + 0.75L) /
                             ((long double)n + 0.5L));
        long double pp = 0.0L;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1.0L, p1 = z;
            for (size_t k = 2; k <= n; ++k) {
                long double p2 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p0) / (long double)k;
                p0 = p1; p1 = p2;
            }
            if (n == 1) { p1 = z; p0 = 1.0L; }
            pp = (long double)n * (z * p1 - p0) / (z * z - 1.0L);
            long double dz = p1 / pp;
            z -= dz;
            if (fabsl(dz) < 1e-19L) break;
        }
        long double p0 = 1.0L, p1 = z;
        for (size_t k = 2; k <= n; ++k) {
            long double p2 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p0) / (long double)k;
            p0 = p1; p1 = p2;
        }
        pp = (long double)n * (z * p1 - p0) / (z * z - 1.0L);
        const long double ww = 2.0L / ((1.0L - z * z) * pp * pp);
        x[i] = -(double)z;
        x[n - 1 - i] = (double)z;
        w[i] = (double)ww;
        w[n - 1 - i] = (double)ww;
    }
    if (n & 1) x[n / 2] = 0.0;
    return ORC_OK;
}

double orc_legendre_pl(int l, double x)
{
    if (l == 0) return 1.0;
    if (l == 1) return x;
    double p0 = 1.0, p1 = x;
    for (int k = 2; k <= l; ++k) {
        const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

double orc_sph_jl(int l, double x)
{
    if (l < 0) return NAN;
    const double ax = fabs(x);
    if (ax == 0.0) return l == 0 ? 1.0 : 0.0;
    if (ax < 0.5) {
        /* x^l/(2l+1)!! * sum_k (-x^2/2)^k / (k! (2l+3)(2l+5)...(2l+2k+1)) */
        double pre = 1.0;
        for (int k = 1; k <= l; ++k) pre *= x / (2.0 * k + 1.0);
        const double y = -0.5 * x * x;
        double term = 1.0, sum = 1.0;
        for (int k = 1; k < 30; ++k) {
            term *= y / (k * (2.0 * l + 2.0 * k + 1.0));
            sum += term;
            if (fabs(term) < 1e-18 * fabs(sum)) break;
        }
        return pre * sum;
    }
    const double s = sin(x), c = cos(x);
    const double j0 = s / x;
    if (l == 0) return j0;
    const double j1 = (s / x - c) / x;
    if (l == 1) return j1;
    if (ax > (double)l) {
        double jm = j0, jc = j1;
        for (int k = 1; k < l; ++k) {
            const double jn = (2.0 * k + 1.0) / x * jc - jm;
            jm = jc;
            jc = jn;
        }
        return jc;
    }
    /* Miller: start well above l; normalise with sum (2k+1) j_k^2 = 1 */
    const int L = l + 30 + (int)ax;
    double fp = 0.0, fc = 1.0, fl = 0.0, norm = 0.0, f0 = 0.0, f1 = 0.0;
    for (int k = L; k >= 0; --k) {
        /* fc holds f_k */
        norm += (2.0 * k + 1.0) * fc * fc;
        if (k == l) fl = fc;
        if (k == 1) f1 = fc;
        if (k == 0) { f0 = fc; break; }
        const double fm = (2.0 * k + 1.0) / x * fc - fp;
        fp = fc;
        fc = fm;
        if (fabs(fc) > 1e250) {   /* rescale to avoid overflow */
            fc *= 1e-250; fp *= 1e-250; fl *= 1e-250; norm *= 1e-250; norm *= 1e-250; f1 *= 1e-250;
        }
    }
    double scale = 1.0 / sqrt(norm);
    /* sign from whichever of j_0, j_1 is better conditioned */
    const double ref = fabs(j0) > fabs(j1) ? j0 : j1;
    const double fref = fabs(j0) > fabs(j1) ? f0 : f1;
    if ((ref < 0) != (fref * scale < 0)) scale = -scale;
    return fl * scale;
}

/* test hook: QAG on closed-form integrands (kind 0: x^9 + cos(x); 1: sin(50 x) e^{-x};
   2: 1/sqrt(x)) so the adaptive bisection can be pinned against exact values */
static double qt0(double x, void *p) { (void)p; return pow(x, 9) + cos(x); }
static double qt1(double x, void *p) { (void)p; return sin(50 * x) * exp(-x); }
static double qt2(double x, void *p) { (void)p; return 1.0 / sqrt(x); }
double orc_qag_test(int kind, double a, double b, double epsrel, int *status)
{
    orc_fn f = kind == 0 ? qt0 : (kind == 1 ? qt1 : qt2);
    return orc_qag(f, NULL, a, b, 0.0, epsrel, 50000, NULL, status);
}
