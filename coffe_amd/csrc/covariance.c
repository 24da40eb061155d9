/*
 * Host side of the 2D FFTLog covariance (C99): plans, the spectrum-independent
 * kernel arrays, the bicubic functionals, the Gaussian covariance assembly and the
 * drop-in two_sph_bessel.  Reference behaviour restated here:
 *   - lngamma_lanczos, g_l, c_window_2d (utils.c:179-258);
 *   - two_sph_bessel (twobessel.c:16-113): eta_m = 2 pi m / (N dlnk), r = 1/k reversed,
 *     kernel (k0 r0)^(-i eta) g_l, the row window on rows N/2 - i and N/2 + 1 + i and
 *     the column window on column N/2 - j, output scale pi / (16 N1 N2 r2^nu2 r1^nu1);
 *   - covariance_integrate_fftlog (covariance.c:357-470): nu1 = nu2 = 1.01, window
 *     width 0.25, k_i = kmin (kmax/kmin)^(i/N), diagonal k^3 P^p / dlnk, bicubic
 *     (natural cubic splines along r1, r2 and r2-of-d/dr1; covariance.c:443-464) at
 *     sep x sep, scaled by (2 l1 + 1)(2 l2 + 1) {2, 1} / pi^2 (covariance.c:754-767);
 *   - the Gaussian covariance of multipoles (covariance.c:796-1060 with
 *     covariance_coefficient 40-57 and coffe_legendre_integral common.c:1370-1394).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "ce_host.h"
#include "ce_internal.h"
#include "coffe_engine.h"
#include "coffe_twobessel.h"

typedef double complex cplx;

#define CV_PI 3.14159265358979323846
#define CV_MAXL 8
#define CV_MAXPAIR (CV_MAXL * (CV_MAXL + 1) / 2)

/* ---- kernel arrays ------------------------------------------------------------ */

static cplx lngamma_lanczos(cplx z)
{
    /* Lanczos approximation, g = 7, 9 coefficients */
    static const double p[] = {0.99999999999980993227684700473478, 676.520368121885098567009190444019,
                               -1259.13921672240287047156078755283, 771.3234287776530788486528258894,
                               -176.61502916214059906584551354,     12.507343278686904814458936853,
                               -0.13857109526572011689554707,       9.984369578019570859563e-6,
                               1.50563273514931155834e-7};
    if (creal(z) < 0.5) return log(CV_PI) - clog(csin(CV_PI * z)) - lngamma_lanczos(1. - z);
    z -= 1;
    cplx x = p[0];
    for (int n = 1; n < 9; n++) x += p[n] / (z + (double)(n));
    const cplx t = z + 7.5;
    return log(2 * CV_PI) / 2. + (z + 0.5) * clog(t) - t + clog(x);
}

/* g_l(z) = 2^z Gamma((l + z)/2) / Gamma((3 + l - z)/2), z = nu + i eta */
static cplx g_l(double l, double nu, double eta)
{
    const cplx z = nu + I * eta;
    return cexp(z * log(2.) + lngamma_lanczos((l + z) / 2.) - lngamma_lanczos((3. + l - z) / 2.));
}

static double cwin(long i, long ncut) { return (double)(i) / ncut - 1. / (2. * CV_PI) * sin(2. * i * CV_PI / ncut); }

/*
 * The two axis arrays of one transform: with H[m] = (k0 r0)^(-i eta_m) g_l(eta_m), m <= N/2,
 *   A[m] (m < N, first axis)  = W_row(m) conj(H[m]) for m <= N/2, W_row(m) H[N - m] above;
 *   B[n] (n <= N/2, second axis) = W_col(n) conj(H[n]),
 * interleaved complex.  The 2D transform multiplies out[m][n] by H2[n] H1[m] (with the
 * conjugate-symmetric continuation for m > N/2) and conjugates (twobessel.c:67-98).
 */
static void axis_arrays(int N, double k0r0, double dlnk, double l, double nu, double width, double *A, double *B)
{
    const long h = N / 2;
    const long ncut = (long)(h * width);
    double *wr = (double *)malloc(sizeof(double) * (size_t)N);
    double *wc = (double *)malloc(sizeof(double) * (size_t)(h + 1));
    cplx *H = (cplx *)malloc(sizeof(cplx) * (size_t)(h + 1));
    if (!wr || !wc || !H) { free(wr); free(wc); free(H); return; }
    for (long i = 0; i < N; ++i) wr[i] = 1.0;
    for (long j = 0; j <= h; ++j) wc[j] = 1.0;
    for (long j = 0; j <= ncut; ++j) wc[h - j] *= cwin(j, ncut);
    for (long i = 0; i <= ncut; ++i) {
        wr[h - i] *= cwin(i, ncut);
        wr[h + 1 + i] *= cwin(i, ncut);
    }
    for (long m = 0; m <= h; ++m) {
        const double eta = 2 * CV_PI / dlnk / N * m;
        H[m] = cexp(-I * eta * clog((cplx)k0r0)) * g_l(l, nu, eta);
    }
    for (long m = 0; m < N; ++m) {
        const cplx a = (m <= h) ? wr[m] * conj(H[m]) : wr[m] * H[N - m];
        A[2 * m] = creal(a);
        A[2 * m + 1] = cimag(a);
    }
    for (long n = 0; n <= h; ++n) {
        const cplx b = wc[n] * conj(H[n]);
        B[2 * n] = creal(b);
        B[2 * n + 1] = cimag(b);
    }
    free(wr); free(wc); free(H);
}

/* the window must fit the row (twobessel.c:76-92); lengths the fused LDS kernels take
   (N <= 8192, N and N/2 of 2, 3, 5) run there, other even lengths up to 2^20 on the
   general device FFT */
static int cov_window_ok(int N, double width)
{
    const long ncut = (long)((N / 2) * width);
    return ncut >= 1 && N / 2 + 1 + ncut <= N - 1;
}

static int cov_lds_ok(int N)
{
    ce_radix_plan rp;
    if (N < 8 || N > 8192 || (N & 1)) return 0;
    return ce_radix_schedule_ex(N, &rp, 1) == COFFE_OK && ce_radix_schedule_ex(N / 2, &rp, 1) == COFFE_OK;
}

static int cov_length_ok(int N, double width)
{
    return N >= 8 && !(N & 1) && N <= (1 << 20) && cov_window_ok(N, width);
}

/* ---- bicubic functionals ------------------------------------------------------- */

/* index i with x[i] <= v < x[i + 1] (i = n - 2 at v = x[n - 1]) */
static int bsearch_lo(const double *x, int n, double v)
{
    int lo = 0, hi = n - 1;
    while (hi > lo + 1) {
        const int m = (lo + hi) / 2;
        if (x[m] > v) hi = m;
        else lo = m;
    }
    return lo;
}

/*
 * Weights w[0..n) of the natural cubic spline's first derivative at node i as a linear
 * functional of the node values.  With c the half second derivatives (T c = g, T the
 * symmetric tridiagonal spline matrix, g_q = 3 (dy_{q+1}/h_{q+1} - dy_q/h_q)), the
 * derivative is alpha . y + beta . c, so w = alpha + E^T T^-1 beta (T is symmetric).
 */
static void spline_deriv_weights(int n, const double *x, int i, double *w, double *work)
{
    const int S = n - 2;
    double *diag = work, *gam = work + n, *beta = work + 2 * n;
    memset(w, 0, sizeof(double) * (size_t)n);
    memset(beta, 0, sizeof(double) * (size_t)n);
    if (i <= n - 2) {
        const double h = x[i + 1] - x[i];
        w[i + 1] += 1.0 / h;
        w[i] -= 1.0 / h;
        /* - h (c_{i+1} + 2 c_i) / 3 */
        if (i + 1 <= n - 2) beta[i] += -h / 3.0;
        if (i >= 1) beta[i - 1] += -2.0 * h / 3.0;
    } else {
        const double h = x[n - 1] - x[n - 2];
        w[n - 1] += 1.0 / h;
        w[n - 2] -= 1.0 / h;
        /* + h (2 c_{n-1} + c_{n-2}) / 3, c_{n-1} = 0 */
        beta[n - 3] += h / 3.0;
    }
    /* gamma = T^-1 beta: LDL^T forward/backward sweeps */
    for (int q = 0; q < S; ++q) {
        const double hq = x[q + 1] - x[q], hq1 = x[q + 2] - x[q + 1];
        diag[q] = 2.0 * (hq + hq1);
    }
    /* forward elimination; the coupling of unknowns q-1 and q is h_q */
    double *z = gam;
    z[0] = beta[0];
    for (int q = 1; q < S; ++q) {
        const double off_prev = x[q + 1] - x[q];
        const double l = off_prev / diag[q - 1];
        diag[q] -= l * off_prev;
        z[q] = beta[q] - l * z[q - 1];
    }
    z[S - 1] /= diag[S - 1];
    for (int q = S - 2; q >= 0; --q) {
        const double off = x[q + 2] - x[q + 1];   /* coupling (q, q+1) = h_{q+1} */
        z[q] = (z[q] - off * z[q + 1]) / diag[q];
    }
    for (int q = 0; q < S; ++q) {
        const double hq = x[q + 1] - x[q], hq1 = x[q + 2] - x[q + 1];
        const double gq = 3.0 * z[q];
        w[q + 2] += gq / hq1;
        w[q + 1] -= gq * (1.0 / hq1 + 1.0 / hq);
        w[q] += gq / hq;
    }
}

typedef struct {
    int R;
    int *lo, *cnt, *off;   /* node window of functional m and its weight offset */
    double *w;
} cv_functionals;

static void functionals_free(cv_functionals *f)
{
    free(f->lo); free(f->cnt); free(f->off); free(f->w);
    memset(f, 0, sizeof(*f));
}

/* the bicubic's dependence on one axis at the query points v[0..R): Hermite value and
   derivative weights of the cell, derivatives as spline functionals; entries below
   1e-20 of the largest weight are dropped (they decay as (2 - sqrt 3)^distance) */
static int build_functionals(int n, const double *x, int R, const double *v, cv_functionals *f)
{
    memset(f, 0, sizeof(*f));
    f->R = R;
    f->lo = (int *)malloc(sizeof(int) * (size_t)R);
    f->cnt = (int *)malloc(sizeof(int) * (size_t)R);
    f->off = (int *)malloc(sizeof(int) * (size_t)R);
    double *u = (double *)malloc(sizeof(double) * (size_t)n);
    double *w0 = (double *)malloc(sizeof(double) * (size_t)n);
    double *w1 = (double *)malloc(sizeof(double) * (size_t)n);
    double *work = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    size_t cap = 0, used = 0;
    int st = COFFE_OK;
    if (!f->lo || !f->cnt || !f->off || !u || !w0 || !w1 || !work) st = COFFE_ENOMEM;
    for (int m = 0; m < R && st == COFFE_OK; ++m) {
        if (!(v[m] >= x[0] && v[m] <= x[n - 1])) { st = COFFE_EDOMAIN; break; }
        const int i = bsearch_lo(x, n, v[m]);
        const double dx = x[i + 1] - x[i];
        const double t = (v[m] - x[i]) / dx;
        const double h0 = (1 + 2 * t) * (1 - t) * (1 - t), h1 = t * t * (3 - 2 * t);
        const double h2 = t * (1 - t) * (1 - t), h3 = t * t * (t - 1);
        spline_deriv_weights(n, x, i, w0, work);
        spline_deriv_weights(n, x, i + 1, w1, work);
        double big = 0.0;
        for (int j = 0; j < n; ++j) {
            u[j] = dx * (h2 * w0[j] + h3 * w1[j]);
            if (j == i) u[j] += h0;
            if (j == i + 1) u[j] += h1;
            if (fabs(u[j]) > big) big = fabs(u[j]);
        }
        int lo = i, hi = i + 1;
        for (int j = 0; j < n; ++j)
            if (fabs(u[j]) > 1e-20 * big) { lo = j < lo ? j : lo; hi = j > hi ? j : hi; }
        const int cnt = hi - lo + 1;
        if (used + (size_t)cnt > cap) {
            cap = 2 * (used + (size_t)cnt) + 64;
            double *nw = (double *)realloc(f->w, sizeof(double) * cap);
            if (!nw) { st = COFFE_ENOMEM; break; }
            f->w = nw;
        }
        f->lo[m] = lo;
        f->cnt[m] = cnt;
        f->off[m] = (int)used;
        memcpy(f->w + used, u + lo, sizeof(double) * (size_t)cnt);
        used += (size_t)cnt;
    }
    free(u); free(w0); free(w1); free(work);
    if (st != COFFE_OK) functionals_free(f);
    f->R = R;
    (void)used;
    return st;
}

/* ---- plans ---------------------------------------------------------------------- */

struct coffe_cov_plan {
    coffe_engine *e;
    int N, M, nl, npair, R, nrows, nw;
    int ls[CV_MAXL];
    int pl1[CV_MAXPAIR], pl2[CV_MAXPAIR];
    double kmin, kmax, step;
    ce_radix_plan rpN, rpM;
    ce_anyfft *fft;    /* the general path's length-N transform (lengths outside the LDS kernels) */
    double *h_sep;
    /* device: spectrum-independent */
    double *d_tw, *d_A, *d_ahat, *d_B, *d_kg, *d_k3, *d_kn1, *d_kn2, *d_prn1, *d_prn2, *d_fac, *d_fw, *d_sep;
    double *d_fws;     /* the column weights times pi / (16 N^2 r_j^nu2) (the row transform's scale) */
    int *d_rows, *d_pl1, *d_pl2, *d_flo, *d_fcnt, *d_foff, *d_upos, *d_status, *d_ls, *d_pair_of;
    int *d_uelo, *d_uehi;   /* monotone envelopes of the windows in row order (adjoint) */
    /* per call, grown on demand */
    int spec_cap;
    int64_t *d_in_off;
    int32_t *d_n_in;
    double *d_d, *d_W;
    /* adjoint workspaces (coffe_cov_integrals_vjp), allocated on first use */
    int adj_cap;
    double *d_Wbar, *d_dbar;
};

void coffe_cov_plan_destroy(coffe_cov_plan *p)
{
    if (!p) return;
    free(p->h_sep);
    double *dd[] = {p->d_tw, p->d_A, p->d_ahat, p->d_B, p->d_kg, p->d_k3, p->d_kn1, p->d_kn2, p->d_prn1,
                    p->d_prn2, p->d_fac, p->d_fw, p->d_fws, p->d_sep, p->d_d, p->d_W};
    for (size_t i = 0; i < sizeof(dd) / sizeof(dd[0]); ++i) hipFree(dd[i]);
    int *di[] = {p->d_rows, p->d_pl1, p->d_pl2, p->d_flo, p->d_fcnt, p->d_foff, p->d_upos, p->d_uelo, p->d_uehi, p->d_status,
                 p->d_ls, p->d_pair_of};
    for (size_t i = 0; i < sizeof(di) / sizeof(di[0]); ++i) hipFree(di[i]);
    hipFree(p->d_in_off);
    hipFree(p->d_n_in);
    hipFree(p->d_Wbar);
    hipFree(p->d_dbar);
    ce_anyfft_destroy(p->fft);
    free(p);
}

static int upload(void **d, const void *h, size_t bytes, hipStream_t s)
{
    int st = ce_dev_alloc(d, bytes);
    if (st == COFFE_OK) st = ce_hip_ok(hipMemcpyAsync(*d, h, bytes, hipMemcpyHostToDevice, s));
    return st;
}

int coffe_cov_plan_create(coffe_engine *e, int N, double kmin, double kmax, int nl, const int *ls, int R,
                          const double *sep, coffe_cov_plan **out)
{
    const double nu = 1.01, width = 0.25;
    if (!e || !out || !ls || !sep || nl <= 0 || nl > CV_MAXL || R <= 0) return COFFE_EINVAL;
    if (!(kmin > 0.0) || !(kmax > kmin)) return COFFE_EDOMAIN;
    if (!cov_length_ok(N, width)) return COFFE_EUNSUPPORTED;
    for (int i = 0; i < nl; ++i)
        if (ls[i] < 0 || ls[i] > 40) return COFFE_EINVAL;
    hipSetDevice(e->device);
    hipStream_t s = e->stream;
    coffe_cov_plan *p = (coffe_cov_plan *)calloc(1, sizeof(*p));
    if (!p) return COFFE_ENOMEM;
    p->e = e; p->N = N; p->M = N / 2; p->nl = nl; p->R = R; p->kmin = kmin; p->kmax = kmax;
    memcpy(p->ls, ls, sizeof(int) * (size_t)nl);
    p->step = log(kmax / kmin) / N;
    int st = COFFE_OK;
    if (cov_lds_ok(N)) {
        st = ce_radix_schedule_ex(N, &p->rpN, 1);
        if (st == COFFE_OK) st = ce_radix_schedule_ex(N / 2, &p->rpM, 1);
    } else {
        st = ce_anyfft_create(N, &p->fft);
    }
    int pair_of[CV_MAXL * CV_MAXL];
    double fac[2 * CV_MAXPAIR];
    p->npair = 0;
    for (int i = 0; i < nl; ++i)
        for (int j = i; j < nl; ++j) {
            p->pl1[p->npair] = i;
            p->pl2[p->npair] = j;
            pair_of[i * nl + j] = pair_of[j * nl + i] = p->npair;
            p->npair++;
        }
    for (int q = 0; q < p->npair; ++q) {
        const int l1 = ls[p->pl1[q]], l2 = ls[p->pl2[q]];
        fac[q] = (2 * l1 + 1) * (2 * l2 + 1) * 2. / CV_PI / CV_PI;
        fac[p->npair + q] = (2 * l1 + 1) * (2 * l2 + 1) / CV_PI / CV_PI;
    }
    /* grids (covariance.c:389-403, twobessel.c:42-45) */
    double *k = (double *)malloc(sizeof(double) * (size_t)N);
    double *r = (double *)malloc(sizeof(double) * (size_t)N);
    double *k3 = (double *)malloc(sizeof(double) * (size_t)N);
    double *kn = (double *)malloc(sizeof(double) * (size_t)N);
    double *prn = (double *)malloc(sizeof(double) * (size_t)N);
    double *tw = (double *)malloc(sizeof(double) * 2 * (size_t)N);
    double *A = (double *)malloc(sizeof(double) * 2 * (size_t)N * (size_t)nl);
    double *B = (double *)malloc(sizeof(double) * 2 * (size_t)(N / 2 + 1) * (size_t)nl);
    p->h_sep = (double *)malloc(sizeof(double) * (size_t)R);
    cv_functionals f;
    memset(&f, 0, sizeof(f));
    int *mark = (int *)calloc((size_t)N, sizeof(int));
    int *pos = (int *)malloc(sizeof(int) * (size_t)N);
    int *rows = (int *)malloc(sizeof(int) * (size_t)N);
    int *upos = (int *)malloc(sizeof(int) * (size_t)R);
    if (st == COFFE_OK && (!k || !r || !k3 || !kn || !prn || !tw || !A || !B || !p->h_sep || !mark || !pos || !rows ||
                           !upos))
        st = COFFE_ENOMEM;
    if (st == COFFE_OK) {
        memcpy(p->h_sep, sep, sizeof(double) * (size_t)R);
        for (int i = 0; i < N; ++i) {
            k[i] = kmin * pow(kmax / kmin, (double)i / N);
            k3[i] = pow(k[i], 3);
            kn[i] = pow(k[i], nu);
        }
        for (int i = 0; i < N; ++i) r[i] = 1. / k[N - 1 - i];
        for (int i = 0; i < N; ++i) prn[i] = pow(r[i], nu);
        ce_twiddles(N, tw);
        const double dlnk = log(k[1] / k[0]);
        for (int i = 0; i < nl; ++i)
            axis_arrays(N, k[0] * r[0], dlnk, (double)ls[i], nu, width, A + 2 * (size_t)N * i,
                        B + 2 * (size_t)(N / 2 + 1) * i);
        st = build_functionals(N, r, R, sep, &f);
    }
    if (st == COFFE_OK) {
        /* rows the functionals touch, in increasing order */
        for (int m = 0; m < R; ++m)
            for (int j = f.lo[m]; j < f.lo[m] + f.cnt[m]; ++j) mark[j] = 1;
        p->nrows = 0;
        for (int j = 0; j < N; ++j)
            if (mark[j]) { pos[j] = p->nrows; rows[p->nrows++] = j; }
        for (int m = 0; m < R; ++m) upos[m] = pos[f.lo[m]];
        p->nw = f.off[R - 1] + f.cnt[R - 1];
    }
    const size_t nb = sizeof(double) * (size_t)N;
    if (st == COFFE_OK) st = upload((void **)&p->d_tw, tw, 2 * nb, s);
    if (st == COFFE_OK) st = upload((void **)&p->d_A, A, 2 * nb * (size_t)nl, s);
    if (st == COFFE_OK) st = ce_dev_alloc((void **)&p->d_ahat, 2 * nb * (size_t)nl);
    if (st == COFFE_OK) st = upload((void **)&p->d_B, B, sizeof(double) * 2 * (size_t)(N / 2 + 1) * (size_t)nl, s);
    if (st == COFFE_OK) st = upload((void **)&p->d_kg, k, nb, s);
    if (st == COFFE_OK) st = upload((void **)&p->d_k3, k3, nb, s);
    if (st == COFFE_OK) st = upload((void **)&p->d_kn1, kn, nb, s);
    if (st == COFFE_OK) st = upload((void **)&p->d_kn2, kn, nb, s);
    if (st == COFFE_OK) st = upload((void **)&p->d_prn1, prn, nb, s);
    if (st == COFFE_OK) st = upload((void **)&p->d_prn2, prn, nb, s);
    if (st == COFFE_OK) st = upload((void **)&p->d_fac, fac, sizeof(double) * 2 * (size_t)p->npair, s);
    if (st == COFFE_OK) st = upload((void **)&p->d_fw, f.w, sizeof(double) * (size_t)p->nw, s);
    if (st == COFFE_OK) {
        double *fws = (double *)malloc(sizeof(double) * (size_t)p->nw);
        if (!fws) st = COFFE_ENOMEM;
        for (int m = 0; m < R && st == COFFE_OK; ++m)
            for (int q = 0; q < f.cnt[m]; ++q)
                fws[f.off[m] + q] = f.w[f.off[m] + q] * (M_PI / (16.0 * (double)N * (double)N * prn[f.lo[m] + q]));
        if (st == COFFE_OK) st = upload((void **)&p->d_fws, fws, sizeof(double) * (size_t)p->nw, s);
        free(fws);
    }
    if (st == COFFE_OK) st = upload((void **)&p->d_sep, sep, sizeof(double) * (size_t)R, s);
    if (st == COFFE_OK) st = upload((void **)&p->d_rows, rows, sizeof(int) * (size_t)p->nrows, s);
    if (st == COFFE_OK) st = upload((void **)&p->d_pl1, p->pl1, sizeof(int) * (size_t)p->npair, s);
    if (st == COFFE_OK) st = upload((void **)&p->d_pl2, p->pl2, sizeof(int) * (size_t)p->npair, s);
    if (st == COFFE_OK) st = upload((void **)&p->d_flo, f.lo, sizeof(int) * (size_t)R, s);
    if (st == COFFE_OK) st = upload((void **)&p->d_fcnt, f.cnt, sizeof(int) * (size_t)R, s);
    if (st == COFFE_OK) st = upload((void **)&p->d_foff, f.off, sizeof(int) * (size_t)R, s);
    if (st == COFFE_OK) st = upload((void **)&p->d_upos, upos, sizeof(int) * (size_t)R, s);
    if (st == COFFE_OK) {
        /* envelopes: start min over m' >= m, end max over m' <= m (both nondecreasing) */
        int *env = (int *)malloc(sizeof(int) * 2 * (size_t)R);
        if (!env) st = COFFE_ENOMEM;
        for (int m = R - 1; m >= 0 && st == COFFE_OK; --m)
            env[m] = (m == R - 1 || upos[m] < env[m + 1]) ? upos[m] : env[m + 1];
        for (int m = 0; m < R && st == COFFE_OK; ++m) {
            const int e = upos[m] + f.cnt[m];
            env[R + m] = (m == 0 || e > env[R + m - 1]) ? e : env[R + m - 1];
        }
        if (st == COFFE_OK) st = upload((void **)&p->d_uelo, env, sizeof(int) * (size_t)R, s);
        if (st == COFFE_OK) st = upload((void **)&p->d_uehi, env + R, sizeof(int) * (size_t)R, s);
        free(env);
    }
    if (st == COFFE_OK) st = upload((void **)&p->d_ls, ls, sizeof(int) * (size_t)nl, s);
    if (st == COFFE_OK) st = upload((void **)&p->d_pair_of, pair_of, sizeof(int) * (size_t)nl * (size_t)nl, s);
    if (st == COFFE_OK) st = ce_dev_alloc((void **)&p->d_status, sizeof(int));
    if (st == COFFE_OK) {
        if (p->fft) st = ce_general_cov_ahat(p->fft, s, nl, N, p->d_A, p->d_ahat);
        else if (ce_launch_cov_ahat(s, nl, N, &p->rpN, p->d_tw, p->d_A, p->d_ahat)) st = COFFE_EDEVICE;
    }
    if (st == COFFE_OK) st = ce_hip_ok(hipStreamSynchronize(s));
    free(k); free(r); free(k3); free(kn); free(prn); free(tw); free(A); free(B);
    free(mark); free(pos); free(rows); free(upos);
    functionals_free(&f);
    if (st != COFFE_OK) { coffe_cov_plan_destroy(p); return st; }
    *out = p;
    return COFFE_OK;
}

int coffe_cov_plan_rows(const coffe_cov_plan *p) { return p ? p->nrows : 0; }
int coffe_cov_plan_pairs(const coffe_cov_plan *p) { return p ? p->npair : 0; }

static int grow(coffe_cov_plan *p, int nspec)
{
    if (nspec <= p->spec_cap) return COFFE_OK;
    hipFree(p->d_in_off); hipFree(p->d_n_in); hipFree(p->d_d); hipFree(p->d_W);
    p->d_in_off = NULL; p->d_n_in = NULL; p->d_d = NULL; p->d_W = NULL; p->spec_cap = 0;
    int st = ce_dev_alloc((void **)&p->d_in_off, sizeof(int64_t) * (size_t)nspec);
    if (st == COFFE_OK) st = ce_dev_alloc((void **)&p->d_n_in, sizeof(int32_t) * (size_t)nspec);
    if (st == COFFE_OK) st = ce_dev_alloc((void **)&p->d_d, sizeof(double) * 2 * (size_t)p->N * (size_t)nspec);
    if (st == COFFE_OK)
        st = ce_dev_alloc((void **)&p->d_W, sizeof(double) * 2 * (size_t)nspec * (size_t)p->npair *
                                                (size_t)p->nrows * (size_t)p->R);
    if (st == COFFE_OK) p->spec_cap = nspec;
    return st;
}

int coffe_cov_integrals(coffe_cov_plan *p, int nspec, const int64_t *n_in, const double *d_k, const double *d_pk,
                        double *d_out, void *stream)
{
    if (!p || nspec <= 0 || !n_in || !d_k || !d_pk || !d_out) return COFFE_EINVAL;
    hipSetDevice(p->e->device);
    hipStream_t s = stream ? (hipStream_t)stream : p->e->stream;
    int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (size_t)nspec);
    int32_t *cnt = (int32_t *)malloc(sizeof(int32_t) * (size_t)nspec);
    if (!off || !cnt) { free(off); free(cnt); return COFFE_ENOMEM; }
    int64_t o = 0;
    int st = COFFE_OK;
    for (int i = 0; i < nspec; ++i) {
        if (n_in[i] < 5 || n_in[i] > INT32_MAX) st = COFFE_EINVAL;
        off[i] = o;
        cnt[i] = (int32_t)n_in[i];
        o += n_in[i];
    }
    if (st == COFFE_OK) st = grow(p, nspec);
    if (st == COFFE_OK) st = ce_hip_ok(hipMemcpyAsync(p->d_in_off, off, sizeof(int64_t) * (size_t)nspec,
                                                      hipMemcpyHostToDevice, s));
    if (st == COFFE_OK) st = ce_hip_ok(hipMemcpyAsync(p->d_n_in, cnt, sizeof(int32_t) * (size_t)nspec,
                                                      hipMemcpyHostToDevice, s));
    if (st == COFFE_OK && ce_launch_cov_diag(s, nspec, p->N, p->d_in_off, p->d_n_in, d_k, d_pk, p->d_kg, p->d_k3,
                                             p->d_kn1, p->d_kn2, p->step, p->d_d, p->d_status))
        st = COFFE_EDEVICE;
    if (st == COFFE_OK && p->fft)
        st = ce_general_cov_rows(p->fft, s, 2 * nspec, p->N, p->d_ahat, p->d_B, p->d_d, p->npair, p->d_pl1, p->d_pl2,
                                 p->nrows, p->d_rows, p->d_prn1, p->R, p->d_flo, p->d_fcnt, p->d_foff, p->d_fws,
                                 p->d_W);
    else if (st == COFFE_OK &&
             ce_launch_cov_rows(s, 2 * nspec, p->N, &p->rpN, &p->rpM, p->d_tw, p->d_ahat, p->d_B, p->d_d, p->npair,
                                p->d_pl1, p->d_pl2, p->nrows, p->d_rows, p->d_prn1, p->d_prn2, p->R, p->d_flo,
                                p->d_fcnt, p->d_foff, p->d_fws, p->d_W))
        st = COFFE_EDEVICE;
    if (st == COFFE_OK && ce_launch_cov_cols(s, nspec, p->npair, p->nrows, p->R, p->d_upos, p->d_fcnt, p->d_foff,
                                             p->d_fw, p->d_fac, p->d_W, d_out))
        st = COFFE_EDEVICE;
    int h_status = 0;
    if (st == COFFE_OK) st = ce_hip_ok(hipMemcpyAsync(&h_status, p->d_status, sizeof(int), hipMemcpyDeviceToHost, s));
    if (st == COFFE_OK) st = ce_hip_ok(hipStreamSynchronize(s));
    if (st == COFFE_OK && h_status) st = COFFE_EDOMAIN;   /* P^p not finite on [kmin, kmax] */
    free(off); free(cnt);
    return st;
}

/* ---- adjoint of the integrals ------------------------------------------------------- */

int coffe_cov_integrals_vjp(coffe_cov_plan *p, int nspec, const int64_t *n_in, const double *d_k, const double *d_pk,
                            const double *d_gint, double *d_gpk, void *stream)
{
    if (!p || nspec <= 0 || !n_in || !d_k || !d_pk || !d_gint || !d_gpk) return COFFE_EINVAL;
    if (p->fft) return COFFE_EUNSUPPORTED;   /* the adjoint runs on the fused LDS kernels only */
    hipSetDevice(p->e->device);
    hipStream_t s = stream ? (hipStream_t)stream : p->e->stream;
    int st = COFFE_OK;
    int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (size_t)nspec);
    int32_t *cnt = (int32_t *)malloc(sizeof(int32_t) * (size_t)nspec);
    if (!off || !cnt) { free(off); free(cnt); return COFFE_ENOMEM; }
    int64_t o = 0;
    for (int i = 0; i < nspec; ++i) {
        if (n_in[i] < 5 || n_in[i] > INT32_MAX) st = COFFE_EINVAL;
        off[i] = o;
        cnt[i] = (int32_t)n_in[i];
        o += n_in[i];
    }
    if (st == COFFE_OK) st = grow(p, nspec);
    if (st == COFFE_OK && nspec > p->adj_cap) {
        hipFree(p->d_Wbar); hipFree(p->d_dbar);
        p->d_Wbar = NULL; p->d_dbar = NULL; p->adj_cap = 0;
        st = ce_dev_alloc((void **)&p->d_Wbar, sizeof(double) * 2 * (size_t)nspec * (size_t)p->npair *
                                                   (size_t)p->nrows * (size_t)p->R);
        if (st == COFFE_OK) st = ce_dev_alloc((void **)&p->d_dbar, sizeof(double) * 2 * (size_t)p->N * (size_t)nspec);
        if (st == COFFE_OK) p->adj_cap = nspec;
    }
    if (st == COFFE_OK) st = ce_hip_ok(hipMemcpyAsync(p->d_in_off, off, sizeof(int64_t) * (size_t)nspec,
                                                      hipMemcpyHostToDevice, s));
    if (st == COFFE_OK) st = ce_hip_ok(hipMemcpyAsync(p->d_n_in, cnt, sizeof(int32_t) * (size_t)nspec,
                                                      hipMemcpyHostToDevice, s));
    if (st == COFFE_OK &&
        ce_launch_cov_vjp(s, nspec, p->N, &p->rpN, &p->rpM, p->d_tw, p->d_ahat, p->d_B, p->npair, p->d_pl1, p->d_pl2,
                          p->nrows, p->d_rows, p->d_prn1, p->R, p->d_flo, p->d_fcnt, p->d_foff, p->d_fws, p->d_upos,
                          p->d_uelo, p->d_uehi, p->d_fw, p->d_fac, d_gint, p->d_Wbar, p->d_dbar, p->d_in_off, p->d_n_in, d_k, d_pk,
                          p->d_kg, p->d_k3, p->d_kn1, p->d_kn2, p->step, d_gpk, o))
        st = COFFE_EDEVICE;
    if (st == COFFE_OK) st = ce_hip_ok(hipStreamSynchronize(s));
    free(off); free(cnt);
    return st;
}

/* ---- assembly --------------------------------------------------------------------- */

static double lfact(int n) { return lgamma((double)n + 1.0); }

/* Wigner 3j (a b c; 0 0 0) in closed form */
static double threej000(int a, int b, int c)
{
    const int J = a + b + c;
    if (J % 2) return 0.0;
    if (c > a + b || c < abs(a - b)) return 0.0;
    const int g = J / 2;
    const double lr = 0.5 * (lfact(J - 2 * a) + lfact(J - 2 * b) + lfact(J - 2 * c) - lfact(J + 1)) + lfact(g) -
                      lfact(g - a) - lfact(g - b) - lfact(g - c);
    return ((g % 2) ? -1.0 : 1.0) * exp(lr);
}

/* int_{-1}^{1} P_n P_m P_a P_b (common.c:1370-1394) */
static double legendre_integral(int n, int m, int a, int b)
{
    if ((n + m + a + b) % 2 != 0) return 0;
    const int lo = abs(n - m) > abs(a - b) ? abs(n - m) : abs(a - b);
    const int hi = n + m < a + b ? n + m : a + b;
    double r = 0;
    for (int l = lo; l <= hi; ++l)
        r += (2. * l + 1) * pow(threej000(n, m, l), 2) * pow(threej000(a, b, l), 2);
    return 2 * r;
}


static int sgnp(int m) { return (abs(m) % 2 == 0) ? 1 : -1; }

static int kdelta(int i, int j) { return i == j; }

/* a value and its derivatives by the four galaxy biases of a bin and f */
typedef struct { double v, d[5]; } cv_dual;

static cv_dual dl_var(double v, int slot)
{
    cv_dual r;
    memset(&r, 0, sizeof(r));
    r.v = v;
    if (slot >= 0) r.d[slot] = 1.0;
    return r;
}

static cv_dual dl_lin(double a, cv_dual x, double b, cv_dual y)   /* a x + b y */
{
    cv_dual r;
    r.v = a * x.v + b * y.v;
    for (int q = 0; q < 5; ++q) r.d[q] = a * x.d[q] + b * y.d[q];
    return r;
}

static cv_dual dl_mul(cv_dual x, cv_dual y)
{
    cv_dual r;
    r.v = x.v * y.v;
    for (int q = 0; q < 5; ++q) r.d[q] = x.d[q] * y.v + x.v * y.d[q];
    return r;
}

/* the coefficient of populations (pa, pb), covariance_coefficient_pair
   (covariance.c:40-74: biases bias[p - 1]), with derivatives */
static cv_dual cov_pair_dual(const cv_dual *gb, int pa, int pb, cv_dual f, int ell)
{
    const cv_dual b1 = gb[pa - 1], b2 = gb[pb - 1];
    const cv_dual sum = dl_lin(1.0, b1, 1.0, b2);
    const cv_dual ff = dl_mul(f, f);
    switch (ell) {
    case 0: return dl_lin(1.0, dl_lin(1.0, dl_mul(b1, b2), 1.0 / 3., dl_mul(sum, f)), 1.0 / 5., ff);
    case 2: return dl_lin(2.0 / 3., dl_mul(sum, f), 4.0 / 7., ff);
    case 4: return dl_lin(8.0 / 35., ff, 0.0, ff);
    default: return dl_var(0.0, -1);
    }
}

/* populations pops[0..3] (NULL: 1, 1, 1, 1) into p[4]; 0 when one is outside 1..4 */
static int read_pops(const int *pops, int *p)
{
    for (int q = 0; q < 4; ++q) {
        p[q] = pops ? pops[q] : 1;
        if (p[q] < 1 || p[q] > 4) return 0;
    }
    return 1;
}

/* the per-(z, l1, l2) coefficients of covariance.c:866-937 for populations pp, biases
   gb[0..3] and f of one bin and the densities n1 (density1) and n2 (density2):
   *cc of the cosmic term, *cm of the mixed term */
static void cov_coefficients(const int *pp, const cv_dual *gb, cv_dual f, double n1, double n2, int l1, int l2,
                             cv_dual *cc, cv_dual *cm)
{
    const int p1 = pp[0], p2 = pp[1], p3 = pp[2], p4 = pp[3];
    cv_dual c = dl_var(0.0, -1), mm = dl_var(0.0, -1);
    for (int a = 0; a <= 4; a += 2)
        for (int b = 0; b <= 4; b += 2) {
            const cv_dual t = dl_lin(1.0, dl_mul(cov_pair_dual(gb, p1, p3, f, a), cov_pair_dual(gb, p2, p4, f, b)),
                                     (double)sgnp(b),
                                     dl_mul(cov_pair_dual(gb, p1, p4, f, a), cov_pair_dual(gb, p2, p3, f, b)));
            c = dl_lin(1.0, c, legendre_integral(l1, l2, a, b), t);
        }
    for (int a = 0; a <= 4; a += 2) {
        cv_dual t = dl_lin(kdelta(p2, p4) / n2, cov_pair_dual(gb, p1, p3, f, a), kdelta(p1, p3) / n1,
                           cov_pair_dual(gb, p2, p4, f, a));
        t = dl_lin(1.0, t, kdelta(p2, p3) / n2, cov_pair_dual(gb, p1, p4, f, a));
        t = dl_lin(1.0, t, kdelta(p1, p4) / n1, cov_pair_dual(gb, p2, p3, f, a));
        mm = dl_lin(1.0, mm, legendre_integral(l1, l2, a, 0), t);
    }
    *cc = dl_lin(0.25, c, 0.0, c);
    *cm = dl_lin(0.125, mm, 0.0, mm);
}

int coffe_cov_assemble(coffe_cov_plan *p, int nspec, int nz, const int *pops, const double *bias, const double *f,
                       const double *D1, const double *nbar1, const double *nbar2, const double *vol,
                       const double *pix, int poisson, int mixed, int cosmic, int window, const double *d_integ,
                       double *d_out, void *stream)
{
    if (!p || nspec <= 0 || nz <= 0 || !bias || !f || !D1 || !nbar1 || !nbar2 || !vol || !pix || !d_integ || !d_out)
        return COFFE_EINVAL;
    int pp[4];
    if (!read_pops(pops, pp)) return COFFE_EINVAL;
    for (int k = 0; k < nz; ++k)
        if (!(nbar1[k] > 0) || !(nbar2[k] > 0) || !(vol[k] > 0) || !(pix[k] > 0)) return COFFE_EDOMAIN;
    hipSetDevice(p->e->device);
    hipStream_t s = stream ? (hipStream_t)stream : p->e->stream;
    const int nl = p->nl;
    const size_t nzl = (size_t)nz * (size_t)nl * (size_t)nl;
    double *cc = (double *)malloc(sizeof(double) * nzl);
    double *cm = (double *)malloc(sizeof(double) * nzl);
    double *zbuf = (double *)malloc(sizeof(double) * 4 * (size_t)nz);
    if (!cc || !cm || !zbuf) { free(cc); free(cm); free(zbuf); return COFFE_ENOMEM; }
    for (int k = 0; k < nz; ++k) {
        cv_dual gb[4];
        for (int q = 0; q < 4; ++q) gb[q] = dl_var(bias[4 * (size_t)k + q], -1);   /* galaxy biases 1..4 */
        for (int i = 0; i < nl; ++i)
            for (int j = 0; j < nl; ++j) {
                cv_dual c, mm;
                cov_coefficients(pp, gb, dl_var(f[k], -1), nbar1[k], nbar2[k], p->ls[i], p->ls[j], &c, &mm);
                cc[((size_t)k * nl + i) * nl + j] = c.v;
                cm[((size_t)k * nl + i) * nl + j] = mm.v;
            }
        zbuf[k] = D1[k];
        zbuf[nz + k] = 1.0 / (nbar1[k] * nbar2[k] * vol[k]);   /* the Poisson term's scale */
        zbuf[2 * nz + k] = vol[k];
        zbuf[3 * nz + k] = pix[k];
    }
    /* the Poisson term's population deltas (covariance.c:945-967) */
    const int pd_same = kdelta(pp[0], pp[2]) * kdelta(pp[1], pp[3]);
    const int pd_swap = kdelta(pp[0], pp[3]) * kdelta(pp[1], pp[2]);
    double *d_cc = NULL, *d_cm = NULL, *d_z = NULL;
    int st = upload((void **)&d_cc, cc, sizeof(double) * nzl, s);
    if (st == COFFE_OK) st = upload((void **)&d_cm, cm, sizeof(double) * nzl, s);
    if (st == COFFE_OK) st = upload((void **)&d_z, zbuf, sizeof(double) * 4 * (size_t)nz, s);
    if (st == COFFE_OK &&
        ce_launch_cov_assemble(s, nspec, nz, nl, p->R, p->npair, p->d_ls, p->d_pair_of, p->d_sep, d_cc, d_cm, d_z,
                               d_z + nz, d_z + 2 * nz, d_z + 3 * nz, poisson ? pd_same : 0, poisson ? pd_swap : 0,
                               window, mixed, cosmic, d_integ, d_out))
        st = COFFE_EDEVICE;
    /* the coefficient buffers are released after the kernel has read them */
    if (st == COFFE_OK) st = ce_hip_ok(hipStreamSynchronize(s));
    hipFree(d_cc); hipFree(d_cm); hipFree(d_z);
    free(cc); free(cm); free(zbuf);
    return st;
}

/* ---- adjoint of the assembly ------------------------------------------------------- */

int coffe_cov_assemble_vjp(coffe_cov_plan *p, int nspec, int nz, const int *pops, const double *bias,
                           const double *f, const double *D1, const double *nbar1, const double *nbar2,
                           const double *vol, int mixed, int cosmic, const double *d_integ, const double *d_gout,
                           double *d_ginteg, double *out_gbias, double *out_gf, void *stream)
{
    if (!p || nspec <= 0 || nz <= 0 || !bias || !f || !D1 || !nbar1 || !nbar2 || !vol || !d_integ || !d_gout ||
        !d_ginteg)
        return COFFE_EINVAL;
    int pp[4];
    if (!read_pops(pops, pp)) return COFFE_EINVAL;
    for (int k = 0; k < nz; ++k)
        if (!(nbar1[k] > 0) || !(nbar2[k] > 0) || !(vol[k] > 0)) return COFFE_EDOMAIN;
    hipSetDevice(p->e->device);
    hipStream_t s = stream ? (hipStream_t)stream : p->e->stream;
    const int nl = p->nl;
    const size_t nzl = (size_t)nz * (size_t)nl * (size_t)nl;
    double *cc = (double *)malloc(sizeof(double) * nzl);
    double *cm = (double *)malloc(sizeof(double) * nzl);
    cv_dual *dc = (cv_dual *)malloc(sizeof(cv_dual) * 2 * nzl);
    double *zbuf = (double *)malloc(sizeof(double) * 2 * (size_t)nz);
    double *gcoef = (double *)malloc(sizeof(double) * 2 * nzl * (size_t)nspec);
    if (!cc || !cm || !dc || !zbuf || !gcoef) {
        free(cc); free(cm); free(dc); free(zbuf); free(gcoef);
        return COFFE_ENOMEM;
    }
    for (int k = 0; k < nz; ++k) {
        cv_dual gb[4];
        for (int q = 0; q < 4; ++q) gb[q] = dl_var(bias[4 * (size_t)k + q], q);
        const cv_dual gr = dl_var(f[k], 4);
        for (int i = 0; i < nl; ++i)
            for (int j = 0; j < nl; ++j) {
                cv_dual c, mm;
                cov_coefficients(pp, gb, gr, nbar1[k], nbar2[k], p->ls[i], p->ls[j], &c, &mm);
                const size_t zl = ((size_t)k * nl + i) * nl + j;
                cc[zl] = c.v;
                cm[zl] = mm.v;
                dc[2 * zl] = mm;
                dc[2 * zl + 1] = c;
            }
        zbuf[k] = D1[k];
        zbuf[nz + k] = vol[k];
    }
    double *d_cc = NULL, *d_cm = NULL, *d_z = NULL, *d_gc = NULL;
    int st = upload((void **)&d_cc, cc, sizeof(double) * nzl, s);
    if (st == COFFE_OK) st = upload((void **)&d_cm, cm, sizeof(double) * nzl, s);
    if (st == COFFE_OK) st = upload((void **)&d_z, zbuf, sizeof(double) * 2 * (size_t)nz, s);
    if (st == COFFE_OK) st = ce_dev_alloc((void **)&d_gc, sizeof(double) * 2 * nzl * (size_t)nspec);
    if (st == COFFE_OK &&
        ce_launch_cov_assemble_vjp(s, nspec, nz, nl, p->R, p->npair, p->d_ls, p->d_pair_of, d_cc, d_cm, d_z, d_z + nz,
                                   mixed, cosmic, d_integ, d_gout, d_ginteg, d_gc))
        st = COFFE_EDEVICE;
    if (st == COFFE_OK)
        st = ce_hip_ok(hipMemcpyAsync(gcoef, d_gc, sizeof(double) * 2 * nzl * (size_t)nspec, hipMemcpyDeviceToHost, s));
    if (st == COFFE_OK) st = ce_hip_ok(hipStreamSynchronize(s));
    if (st == COFFE_OK)
        for (int sp = 0; sp < nspec; ++sp)
            for (int k = 0; k < nz; ++k) {
                double g[5] = {0, 0, 0, 0, 0};
                for (size_t e = 0; e < (size_t)nl * nl; ++e) {
                    const size_t zl = (size_t)k * nl * nl + e;
                    const double *gc = gcoef + 2 * ((size_t)sp * nzl + zl);
                    for (int q = 0; q < 5; ++q) g[q] += gc[0] * dc[2 * zl].d[q] + gc[1] * dc[2 * zl + 1].d[q];
                }
                if (out_gbias)
                    for (int q = 0; q < 4; ++q) out_gbias[((size_t)sp * nz + k) * 4 + q] = g[q];
                if (out_gf) out_gf[(size_t)sp * nz + k] = g[4];
            }
    hipFree(d_cc); hipFree(d_cm); hipFree(d_z); hipFree(d_gc);
    free(cc); free(cm); free(dc); free(zbuf); free(gcoef);
    return st;
}

/* ---- drop-in two_sph_bessel ------------------------------------------------------ */

int two_sph_bessel_status(const double *k1, const double *k2, double **fk1k2, long N1, long N2, const config *cfg,
                          double *r1, double *r2, double **result)
{
    if (!k1 || !k2 || !fk1k2 || !cfg || !r1 || !r2 || !result) return COFFE_EINVAL;
    if ((N1 % 2) || (N2 % 2)) return COFFE_EINVAL;   /* the reference exits on odd lengths */
    if (N1 != N2) return COFFE_EUNSUPPORTED;
    const int N = (int)N1;
    if (!cov_length_ok(N, cfg->c_window_width)) return COFFE_EUNSUPPORTED;
    for (int i = 0; i < N; ++i)
        if (k1[i] != k2[i] || !(k1[i] > 0.0)) return COFFE_EUNSUPPORTED;
    /* the transform is computed for diagonal inputs (the covariance's k^3 P^p / dlnk) */
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            if (i != j && fk1k2[i][j] != 0.0) return COFFE_EUNSUPPORTED;
    coffe_engine *e = coffe_engine_default();
    if (!e) return COFFE_ENODEVICE;
    hipSetDevice(e->device);
    hipStream_t s = e->stream;
    const int M = N / 2;
    for (int i = 0; i < N; ++i) r1[i] = 1. / k1[N - 1 - i];
    for (int i = 0; i < N; ++i) r2[i] = 1. / k2[N - 1 - i];
    ce_radix_plan rpN, rpM;
    ce_anyfft *fft = NULL;
    if (cov_lds_ok(N)) {
        ce_radix_schedule_ex(N, &rpN, 1);
        ce_radix_schedule_ex(M, &rpM, 1);
    } else {
        const int fst = ce_anyfft_create(N, &fft);
        if (fst != COFFE_OK) return fst;
    }
    double *tw = (double *)malloc(sizeof(double) * 2 * (size_t)N);
    double *A = (double *)malloc(sizeof(double) * 2 * (size_t)N);
    double *B = (double *)malloc(sizeof(double) * 2 * (size_t)(M + 1));
    double *dg = (double *)malloc(sizeof(double) * (size_t)N);
    double *p1 = (double *)malloc(sizeof(double) * (size_t)N);
    double *p2 = (double *)malloc(sizeof(double) * (size_t)N);
    int *idx = (int *)malloc(sizeof(int) * (size_t)N);
    int *one = (int *)malloc(sizeof(int) * (size_t)N);
    double *wone = (double *)malloc(sizeof(double) * (size_t)N);
    double *res = (double *)malloc(sizeof(double) * (size_t)N * (size_t)N);
    int st = (tw && A && B && dg && p1 && p2 && idx && one && wone && res) ? COFFE_OK : COFFE_ENOMEM;
    if (st == COFFE_OK) {
        ce_twiddles(N, tw);
        axis_arrays(N, k1[0] * r1[0], log(k1[1] / k1[0]), cfg->l1, cfg->nu1, cfg->c_window_width, A, B);
        double *Bt = (double *)malloc(sizeof(double) * 2 * (size_t)N);
        if (!Bt) st = COFFE_ENOMEM;
        else {
            axis_arrays(N, k2[0] * r2[0], log(k2[1] / k2[0]), cfg->l2, cfg->nu2, cfg->c_window_width, Bt, B);
            free(Bt);
        }
        for (int t = 0; t < N; ++t) {
            dg[t] = fk1k2[t][t] / pow(k1[t], cfg->nu1) / pow(k2[t], cfg->nu2);
            p1[t] = pow(r1[t], cfg->nu1);
            p2[t] = pow(r2[t], cfg->nu2);
            idx[t] = t;
            one[t] = 1;
            wone[t] = M_PI / (16.0 * (double)N * (double)N * p2[t]);   /* the row transform's scale */
        }
    }
    double *d_tw = NULL, *d_A = NULL, *d_ahat = NULL, *d_B = NULL, *d_d = NULL, *d_p1 = NULL, *d_p2 = NULL;
    double *d_w = NULL, *d_W = NULL;
    int *d_idx = NULL, *d_one = NULL, *d_zero = NULL;
    const int zero = 0;
    pthread_mutex_lock(&e->lock);
    if (st == COFFE_OK) st = upload((void **)&d_tw, tw, sizeof(double) * 2 * (size_t)N, s);
    if (st == COFFE_OK) st = upload((void **)&d_A, A, sizeof(double) * 2 * (size_t)N, s);
    if (st == COFFE_OK) st = ce_dev_alloc((void **)&d_ahat, sizeof(double) * 2 * (size_t)N);
    if (st == COFFE_OK) st = upload((void **)&d_B, B, sizeof(double) * 2 * (size_t)(M + 1), s);
    if (st == COFFE_OK) st = upload((void **)&d_d, dg, sizeof(double) * (size_t)N, s);
    if (st == COFFE_OK) st = upload((void **)&d_p1, p1, sizeof(double) * (size_t)N, s);
    if (st == COFFE_OK) st = upload((void **)&d_p2, p2, sizeof(double) * (size_t)N, s);
    if (st == COFFE_OK) st = upload((void **)&d_idx, idx, sizeof(int) * (size_t)N, s);
    if (st == COFFE_OK) st = upload((void **)&d_one, one, sizeof(int) * (size_t)N, s);
    if (st == COFFE_OK) st = upload((void **)&d_w, wone, sizeof(double) * (size_t)N, s);
    if (st == COFFE_OK) st = upload((void **)&d_zero, &zero, sizeof(int), s);
    if (st == COFFE_OK) st = ce_dev_alloc((void **)&d_W, sizeof(double) * (size_t)N * (size_t)N);
    if (st == COFFE_OK) {
        if (fft) st = ce_general_cov_ahat(fft, s, 1, N, d_A, d_ahat);
        else if (ce_launch_cov_ahat(s, 1, N, &rpN, d_tw, d_A, d_ahat)) st = COFFE_EDEVICE;
    }
    /* every row, identity column functionals: W is the N x N result */
    if (st == COFFE_OK && fft)
        st = ce_general_cov_rows(fft, s, 1, N, d_ahat, d_B, d_d, 1, d_zero, d_zero, N, d_idx, d_p1, N, d_idx, d_one,
                                 d_idx, d_w, d_W);
    else if (st == COFFE_OK && ce_launch_cov_rows(s, 1, N, &rpN, &rpM, d_tw, d_ahat, d_B, d_d, 1, d_zero, d_zero, N,
                                                  d_idx, d_p1, d_p2, N, d_idx, d_one, d_idx, d_w, d_W))
        st = COFFE_EDEVICE;
    if (st == COFFE_OK)
        st = ce_hip_ok(hipMemcpyAsync(res, d_W, sizeof(double) * (size_t)N * (size_t)N, hipMemcpyDeviceToHost, s));
    if (st == COFFE_OK) st = ce_hip_ok(hipStreamSynchronize(s));
    pthread_mutex_unlock(&e->lock);
    if (st == COFFE_OK)
        for (int i = 0; i < N; ++i) memcpy(result[i], res + (size_t)i * N, sizeof(double) * (size_t)N);
    double *dd[] = {d_tw, d_A, d_ahat, d_B, d_d, d_p1, d_p2, d_w, d_W};
    for (size_t i = 0; i < sizeof(dd) / sizeof(dd[0]); ++i) hipFree(dd[i]);
    hipFree(d_idx); hipFree(d_one); hipFree(d_zero);
    ce_anyfft_destroy(fft);
    free(tw); free(A); free(B); free(dg); free(p1); free(p2); free(idx); free(one); free(wone); free(res);
    return st;
}

void two_sph_bessel(double *k1, double *k2, double **fk1k2, long N1, long N2, config *cfg, double *r1, double *r2,
                    double **result)
{
    ce_set_last_error(two_sph_bessel_status(k1, k2, fk1k2, N1, N2, cfg, r1, r2, result));
}
