/*
 * Oracle (test infrastructure only): 2D FFTLog and the flat-sky Gaussian
 * covariance of multipoles, restated from
 *   lngamma_lanczos      utils.c:179-199 (Lanczos g = 7, 9 coefficients)
 *   g_l                  utils.c:201-210
 *   c_window_2d          utils.c:224-258
 *   two_sph_bessel       twobessel.c:16-113 (C branch)
 *   covariance_integrate_fftlog   covariance.c:357-470
 *   covariance_coefficient        covariance.c:40-57
 *   coffe_legendre_integral       common.c:1370-1394 (3j symbols in closed form)
 *   coffe_covariance_init (FFTLog method, multipoles output, one population)
 *                                 covariance.c:477-1103
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef double complex cplx;
#define ORC_PI 3.14159265358979323846

static cplx lngamma_lanczos(cplx z)
{
    static const double p[] = {0.99999999999980993227684700473478, 676.520368121885098567009190444019,
                               -1259.13921672240287047156078755283, 771.3234287776530788486528258894,
                               -176.61502916214059906584551354,     12.507343278686904814458936853,
                               -0.13857109526572011689554707,       9.984369578019570859563e-6,
                               1.50563273514931155834e-7};
    if (creal(z) < 0.5) return log(ORC_PI) - clog(csin(ORC_PI * z)) - lngamma_lanczos(1. - z);
    z -= 1;
    cplx x = p[0];
    for (int n = 1; n < 9; n++) x += p[n] / (z + (double)(n));
    cplx t = z + 7.5;
    return log(2 * ORC_PI) / 2. + (z + 0.5) * clog(t) - t + clog(x);
}

void orc_g_l(double l, double nu, size_t n, const double *eta, double *gl)
{
    for (size_t i = 0; i < n; ++i) {
        cplx z = nu + I * eta[i];
        cplx g = cexp(z * log(2.) + lngamma_lanczos((l + z) / 2.) - lngamma_lanczos((3. + l - z) / 2.));
        gl[2 * i] = creal(g);
        gl[2 * i + 1] = cimag(g);
    }
}

static void c_window_2d(cplx *out, double width, long halfN1, long halfN2)
{
    const long N1 = 2 * halfN1;
    const long Ncut1 = (long)(halfN1 * width);
    const long Ncut2 = (long)(halfN2 * width);
    for (long j = 0; j <= Ncut2; j++) {
        const double W = (double)(j) / Ncut2 - 1. / (2. * ORC_PI) * sin(2. * j * ORC_PI / Ncut2);
        for (long i = 0; i < N1; i++) out[i * (halfN2 + 1) + halfN2 - j] *= W;
    }
    for (long i = 0; i <= Ncut1; i++) {
        const double W = (double)(i) / Ncut1 - 1. / (2. * ORC_PI) * sin(2. * i * ORC_PI / Ncut1);
        for (long j = 0; j <= halfN2; j++) {
            out[(halfN1 - i) * (halfN2 + 1) + j] *= W;
            out[(halfN1 + 1 + i) * (halfN2 + 1) + j] *= W;
        }
    }
}

/* c_window_2d on an N1 x (halfN2 + 1) interleaved complex array (the tests' handle) */
void orc_window_2d(double *out, double width, long halfN1, long halfN2)
{
    c_window_2d((cplx *)out, width, halfN1, halfN2);
}

/* fk: N1 x N2 row-major; result: N1 x N2 row-major */
int orc_two_sph_bessel(const double *k1, const double *k2, const double *fk, long N1, long N2, double l1, double l2,
                       double nu1, double nu2, double width, double *r1, double *r2, double *result)
{
    if ((N1 % 2) || (N2 % 2) || N1 < 4 || N2 < 4) return ORC_EINVAL;
    const long h1 = N1 / 2, h2 = N2 / 2;
    const double k10 = k1[0], k20 = k2[0];
    const double dlnk1 = log(k1[1] / k10), dlnk2 = log(k2[1] / k20);
    double *eta_m = (double *)malloc(sizeof(double) * (h1 + 1));
    double *eta_n = (double *)malloc(sizeof(double) * (h2 + 1));
    double *g1 = (double *)malloc(sizeof(double) * 2 * (h1 + 1));
    double *g2 = (double *)malloc(sizeof(double) * 2 * (h2 + 1));
    double *Pb = (double *)malloc(sizeof(double) * N1 * N2);
    cplx *out = (cplx *)malloc(sizeof(cplx) * N1 * (h2 + 1));
    int st = ORC_ENOMEM;
    if (!eta_m || !eta_n || !g1 || !g2 || !Pb || !out) goto done;
    for (long i = 0; i <= h1; i++) eta_m[i] = 2 * ORC_PI / dlnk1 / N1 * i;
    for (long j = 0; j <= h2; j++) eta_n[j] = 2 * ORC_PI / dlnk2 / N2 * j;
    orc_g_l(l1, nu1, (size_t)(h1 + 1), eta_m, g1);
    orc_g_l(l2, nu2, (size_t)(h2 + 1), eta_n, g2);
    for (long i = 0; i < N1; i++) r1[i] = 1. / k1[N1 - 1 - i];
    for (long i = 0; i < N2; i++) r2[i] = 1. / k2[N2 - 1 - i];
    const double r10 = r1[0], r20 = r2[0];
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N1; i++)
        for (long j = 0; j < N2; j++) Pb[i * N2 + j] = fk[i * N2 + j] / pow(k1[i], nu1) / pow(k2[j], nu2);
    st = orc_rfft2_r2c((size_t)N1, (size_t)N2, Pb, (double *)out);
    if (st != ORC_OK) goto done;
    c_window_2d(out, width, h1, h2);
#pragma omp parallel for schedule(static)
    for (long i = 0; i <= h1; i++) {
        for (long j = 0; j <= h2; j++) {
            const long ij = i * (h2 + 1) + j;
            const cplx G2 = g2[2 * j] + I * g2[2 * j + 1];
            const cplx G1 = g1[2 * i] + I * g1[2 * i + 1];
            out[ij] *= cexp(-I * eta_n[j] * clog((cplx)(k20 * r20))) * G2 * cexp(-I * eta_m[i] * clog((cplx)(k10 * r10))) * G1;
            out[ij] = conj(out[ij]);
        }
    }
#pragma omp parallel for schedule(static)
    for (long i = h1 + 1; i < N1; i++) {
        for (long j = 0; j <= h2; j++) {
            const long ij = i * (h2 + 1) + j;
            const cplx G2 = g2[2 * j] + I * g2[2 * j + 1];
            const cplx G1 = g1[2 * (N1 - i)] + I * g1[2 * (N1 - i) + 1];
            out[ij] *= cexp(-I * eta_n[j] * clog((cplx)(k20 * r20))) * G2 *
                       cexp(I * eta_m[N1 - i] * clog((cplx)(k10 * r10))) * conj(G1);
            out[ij] = conj(out[ij]);
        }
    }
    st = orc_rfft2_c2r((size_t)N1, (size_t)N2, (const double *)out, Pb);
    if (st != ORC_OK) goto done;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N1; i++)
        for (long j = 0; j < N2; j++)
            result[i * N2 + j] = Pb[i * N2 + j] * ORC_PI / (16. * N1 * N2 * pow(r2[j], nu2) * pow(r1[i], nu1));
    st = ORC_OK;
done:
    free(eta_m); free(eta_n); free(g1); free(g2); free(Pb); free(out);
    return st;
}

/* covariance.c:357-470: D/G integral on sep x sep, result[R*n + m] at (sep_m, sep_n) */
int orc_covariance_fftlog_integral(size_t nin, const double *kin, const double *pin, double power, int l1, int l2,
                                   size_t N, size_t R, const double *sep, double kmin, double kmax, double *result)
{
    double *akb = (double *)malloc(sizeof(double) * 3 * nin);
    double *k = (double *)malloc(sizeof(double) * N);
    double *fk = (double *)calloc(N * N, sizeof(double));
    double *r1 = (double *)malloc(sizeof(double) * N);
    double *r2 = (double *)malloc(sizeof(double) * N);
    double *res = (double *)malloc(sizeof(double) * N * N);
    double *res2 = (double *)malloc(sizeof(double) * N * N);
    int st = ORC_ENOMEM;
    orc_bicubic *bc = NULL;
    if (!akb || !k || !fk || !r1 || !r2 || !res || !res2) goto done;
    st = orc_akima_init(nin, kin, pin, akb, akb + nin, akb + 2 * nin);
    if (st != ORC_OK) goto done;
    const double step = log(kmax / kmin) / N;
    for (size_t i = 0; i < N; ++i) {
        k[i] = kmin * pow(kmax / kmin, (double)i / N);
        fk[i * N + i] = pow(k[i], 3) * pow(orc_akima_eval(nin, kin, pin, akb, akb + nin, akb + 2 * nin, k[i]), power) / step;
    }
    st = orc_two_sph_bessel(k, k, fk, (long)N, (long)N, l1, l2, 1.01, 1.01, 0.25, r1, r2, res);
    if (st != ORC_OK) goto done;
#pragma omp parallel for schedule(static)
    for (long m = 0; m < (long)N; ++m)
        for (size_t n = 0; n < N; ++n) res2[n * N + (size_t)m] = res[(size_t)m * N + n];
    bc = orc_bicubic_new(N, N, r1, r2, res2);
    if (!bc) { st = ORC_ENOMEM; goto done; }
    for (size_t m = 0; m < R; ++m)
        for (size_t n = 0; n < R; ++n) result[R * n + m] = orc_bicubic_eval(bc, sep[m], sep[n]);
    st = ORC_OK;
done:
    orc_bicubic_free(bc);
    free(akb); free(k); free(fk); free(r1); free(r2); free(res); free(res2);
    return st;
}

static double lfact(int n) { return lgamma((double)n + 1.0); }

double orc_threej000(int a, int b, int c)
{
    const int J = a + b + c;
    if (J % 2) return 0.0;
    if (c > a + b || c < abs(a - b)) return 0.0;
    const int g = J / 2;
    const double lr = 0.5 * (lfact(J - 2 * a) + lfact(J - 2 * b) + lfact(J - 2 * c) - lfact(J + 1)) + lfact(g) -
                      lfact(g - a) - lfact(g - b) - lfact(g - c);
    return ((g % 2) ? -1.0 : 1.0) * exp(lr);
}

double orc_legendre_integral(int n, int m, int a, int b)
{
    if ((n + m + a + b) % 2 != 0) return 0;
    const int lo = abs(n - m) > abs(a - b) ? abs(n - m) : abs(a - b);
    const int hi = n + m < a + b ? n + m : a + b;
    double r = 0;
    for (int l = lo; l <= hi; ++l)
        r += (2. * l + 1) * pow(orc_threej000(n, m, l), 2) * pow(orc_threej000(a, b, l), 2);
    return 2 * r;
}

double orc_covariance_coefficient(double b1, double b2, double f, int ell)
{
    switch (ell) {
    case 0: return b1 * b2 + (b1 + b2) * f / 3 + pow(f, 2) / 5.;
    case 2: return 2 * (b1 + b2) * f / 3. + 4 * pow(f, 2) / 7.;
    case 4: return 8 * pow(f, 2) / 35.;
    default: return 0;
    }
}

static int sgn(int m) { return (abs(m) % 2 == 0) ? 1 : -1; }

static int kdelta(int i, int j) { return i == j ? 1 : 0; }   /* common.c:1401-1407 */

/* covariance_coefficient_pair (covariance.c:60-74): the coefficient of populations
   (p1, p2), biases bias[p - 1] */
static double coefficient_pair(int p1, int p2, const double *bias, double f, int ell)
{
    return orc_covariance_coefficient(bias[p1 - 1], bias[p2 - 1], f, ell);
}

/*
 * Gaussian flat-sky covariance of multipoles, FFTLog method, linear P(k)
 * (redshift-independent integrals scaled by D1^2 and D1^4), covariance.c:790-1060.
 * Populations pops[0..3] (covariance_pop1..4, each 1..4; NULL = 1, 1, 1, 1); per
 * redshift bin k the galaxy biases 1..4 at bias[4 k .. 4 k + 3], growth f[k], D1[k],
 * the densities of the first and second tracer nbar1[k], nbar2[k] (density1/2), the
 * volume vol[k] and the pixel size pix[k].  window selects the bin-averaged Poisson
 * term (covariance_window, covariance.c:931-944).  Output ordered (z, l1, r1, l2, r2).
 * pk_int / pk2_int: L*L blocks of R*R integrals (already scaled), only the l1 <= l2
 * blocks are read; the others are filled by transposition as the reference does
 * (covariance.c:779-788).
 */
int orc_covariance_assemble(size_t nz, size_t nl, const int *ls, size_t R, const double *sep, const int *pops,
                            const double *bias, const double *f, const double *D1, const double *nbar1,
                            const double *nbar2, const double *vol, const double *pix, int poisson, int mixed,
                            int cosmic, int window, const double *pk_int, const double *pk2_int, double *out)
{
    static const int one[4] = {1, 1, 1, 1};
    const int *pp = pops ? pops : one;
    for (int q = 0; q < 4; ++q)
        if (pp[q] < 1 || pp[q] > 4) return ORC_EINVAL;
    const int p1 = pp[0], p2 = pp[1], p3 = pp[2], p4 = pp[3];
    size_t idx = 0;
    const size_t RR = R * R;
    for (size_t k = 0; k < nz; ++k) {
        const double *gb = bias + 4 * k;
        const double gr = f[k];
        const double D = D1[k];
        for (size_t i = 0; i < nl; ++i)
            for (size_t m = 0; m < R; ++m)
                for (size_t j = 0; j < nl; ++j)
                    for (size_t n = 0; n < R; ++n) {
                        const int l1 = ls[i], l2 = ls[j];
                        double cc = 0;
                        for (int a = 0; a <= 4; a += 2)
                            for (int b = 0; b <= 4; b += 2)
                                cc += (coefficient_pair(p1, p3, gb, gr, a) * coefficient_pair(p2, p4, gb, gr, b) +
                                       sgn(b) * coefficient_pair(p1, p4, gb, gr, a) *
                                           coefficient_pair(p2, p3, gb, gr, b)) *
                                      orc_legendre_integral(l1, l2, a, b);
                        cc /= 4.;
                        double cm = 0;
                        for (int a = 0; a <= 4; a += 2)
                            cm += (coefficient_pair(p1, p3, gb, gr, a) * kdelta(p2, p4) / nbar2[k] +
                                   coefficient_pair(p2, p4, gb, gr, a) * kdelta(p1, p3) / nbar1[k] +
                                   coefficient_pair(p1, p4, gb, gr, a) * kdelta(p2, p3) / nbar2[k] +
                                   coefficient_pair(p2, p3, gb, gr, a) * kdelta(p1, p4) / nbar1[k]) *
                                  orc_legendre_integral(l1, l2, a, 0);
                        cm /= 8.;
                        const int dll = l1 == l2, dmn = m == n;
                        double dpk, dpk2;
                        if (i <= j) {
                            dpk = pk_int[(i * nl + j) * RR + R * n + m];
                            dpk2 = pk2_int[(i * nl + j) * RR + R * n + m];
                        } else {
                            dpk = pk_int[(j * nl + i) * RR + R * m + n];
                            dpk2 = pk2_int[(j * nl + i) * RR + R * m + n];
                        }
                        double geom;
                        if (window) {
                            const double lo = sep[m] - pix[k] / 2., hi = sep[m] + pix[k] / 2.;
                            geom = 1. / ((pow(hi, 3) - pow(lo, 3)) / 3.);
                        } else {
                            geom = 1. / sep[n] / sep[m] / pix[k];
                        }
                        const double cov_p =
                            poisson ? (2 * l1 + 1) / 4. / ORC_PI * dll * dmn *
                                          (kdelta(p1, p3) * kdelta(p2, p4) + sgn(l2) * kdelta(p1, p4) * kdelta(p2, p3)) *
                                          geom / nbar1[k] / nbar2[k] / vol[k]
                                    : 0;
                        const double cov_m = mixed ? sgn((l2 - l1) / 2) * sgn(l1 + l2) * D * D * dpk * cm / vol[k] : 0;
                        const double cov_c = cosmic ? sgn((l2 - l1) / 2) * D * D * D * D * dpk2 * cc / vol[k] : 0;
                        out[idx++] = cov_p + cov_m + cov_c;
                    }
    }
    return ORC_OK;
}
