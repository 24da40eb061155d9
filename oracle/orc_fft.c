/*
 * Oracle (test infrastructure only): CPU discrete Fourier transforms.
 *
 * Stands in for FFTW 3.3.9's r2c/c2r (reference call sites twofast.c:272,292,453,465
 * and twobessel.c:60-61,100-102).  Semantics restated from FFTW's documented
 * definitions: forward r2c Y_k = sum_n y_n e^{-2 pi i k n/N}, k = 0..N/2; backward
 * c2r x_n = sum_{k=0}^{N-1} X_k e^{+2 pi i k n/N} with Hermitian completion, no
 * normalisation, imaginary parts of X_0 and X_{N/2} ignored.
 *
 * Algorithm: self-sorting Stockham mixed-radix passes (any factor, hand-written
 * radix 2/3/4/5 butterflies, a direct DFT for other small primes) and Bluestein's
 * chirp-z for lengths with a prime factor > 97.  Twiddles come from long double
 * sin/cos.  Real transforms use the half-length complex transform with the usual
 * split step.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef double complex cplx;

#define ORC_MAXFAC 64

typedef struct orc_fft_plan {
    size_t n;
    int nf;
    int fac[ORC_MAXFAC];
    cplx *tw;           /* e^{-2 pi i k / n}, k < n */
    /* Bluestein */
    int bluestein;
    size_t m;           /* power-of-two convolution length */
    cplx *chirp;        /* e^{-i pi k^2 / n}, k < n */
    cplx *chirp_fft;    /* FFT_m of the conjugate chirp, wrapped */
    struct orc_fft_plan *sub;
} orc_fft_plan;

static cplx twiddle_ld(size_t k, size_t n)
{
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
    return (double)cosl(a) + I * (double)sinl(a);
}

static int factorize(size_t n, int *fac)
{
    int nf = 0;
    static const int pref[] = {4, 2, 3, 5};
    for (int p = 0; p < 4; ++p) {
        while (n % (size_t)pref[p] == 0) { fac[nf++] = pref[p]; n /= (size_t)pref[p]; }
    }
    for (size_t f = 7; n > 1; f += 2) {
        while (n % f == 0) {
            if (nf >= ORC_MAXFAC) return -1;
            fac[nf++] = (int)f;
            n /= f;
        }
    }
    return nf;
}

static void plan_free(orc_fft_plan *p)
{
    if (!p) return;
    free(p->tw);
    free(p->chirp);
    free(p->chirp_fft);
    plan_free(p->sub);
    free(p);
}

static void execute(const orc_fft_plan *p, cplx *data, cplx *work);

static orc_fft_plan *plan_new(size_t n)
{
    orc_fft_plan *p = (orc_fft_plan *)calloc(1, sizeof(*p));
    if (!p) return NULL;
    p->n = n;
    p->nf = factorize(n, p->fac);
    int maxf = 1;
    for (int i = 0; i < p->nf; ++i) if (p->fac[i] > maxf) maxf = p->fac[i];
    if (p->nf < 0 || maxf > 97) {
        p->bluestein = 1;
        size_t m = 1;
        while (m < 2 * n - 1) m <<= 1;
        p->m = m;
        p->chirp = (cplx *)malloc(sizeof(cplx) * n);
        p->chirp_fft = (cplx *)calloc(m, sizeof(cplx));
        p->sub = plan_new(m);
        if (!p->chirp || !p->chirp_fft || !p->sub) { plan_free(p); return NULL; }
        for (size_t k = 0; k < n; ++k) {
            /* k^2 mod 2n keeps the chirp argument exact */
            unsigned long long kk = ((unsigned long long)k * k) % (2ULL * n);
            long double a = -3.14159265358979323846264338327950288L * (long double)kk / (long double)n;
            p->chirp[k] = (double)cosl(a) + I * (double)sinl(a);
        }
        p->chirp_fft[0] = conj(p->chirp[0]);
        for (size_t k = 1; k < n; ++k) {
            p->chirp_fft[k] = conj(p->chirp[k]);
            p->chirp_fft[m - k] = conj(p->chirp[k]);
        }
        cplx *w = (cplx *)malloc(sizeof(cplx) * m);
        if (!w) { plan_free(p); return NULL; }
        execute(p->sub, p->chirp_fft, w);
        free(w);
        return p;
    }
    p->tw = (cplx *)malloc(sizeof(cplx) * (n ? n : 1));
    if (!p->tw) { plan_free(p); return NULL; }
    for (size_t k = 0; k < n; ++k) p->tw[k] = twiddle_ld(k, n);
    return p;
}

/* one Stockham pass: radix R, Ns = product of the radices already applied */
static void stockham_pass(const orc_fft_plan *p, int R, size_t Ns, const cplx *in, cplx *out)
{
    const size_t n = p->n;
    const size_t nr = n / (size_t)R;
    const size_t step = n / (Ns * (size_t)R);   /* twiddle index stride */
    cplx v[97], t[97];
    for (size_t j = 0; j < nr; ++j) {
        const size_t jm = j % Ns;
        for (int r = 0; r < R; ++r) {
            cplx x = in[j + (size_t)r * nr];
            if (r && jm) x *= p->tw[((size_t)r * jm * step) % n];
            v[r] = x;
        }
        switch (R) {
        case 2: {
            cplx a = v[0], b = v[1];
            v[0] = a + b; v[1] = a - b;
        } break;
        case 3: {
            const double s3 = 0.86602540378443864676372317075294;
            cplx a = v[0], b = v[1], c = v[2];
            cplx s = b + c, d = b - c;
            v[0] = a + s;
            cplx m = a - 0.5 * s;
            cplx rot = -I * s3 * d;
            v[1] = m + rot; v[2] = m - rot;
        } break;
        case 4: {
            cplx a = v[0], b = v[1], c = v[2], d = v[3];
            cplx s0 = a + c, d0 = a - c, s1 = b + d, d1 = (b - d) * (-I);
            v[0] = s0 + s1; v[2] = s0 - s1; v[1] = d0 + d1; v[3] = d0 - d1;
        } break;
        case 5: {
            const double c1 = 0.30901699437494742410229341718282;   /* cos(2pi/5) */
            const double c2 = -0.80901699437494742410229341718282;  /* cos(4pi/5) */
            const double s1 = 0.95105651629515357211643933337938;   /* sin(2pi/5) */
            const double s2 = 0.58778525229247312916870595463907;   /* sin(4pi/5) */
            cplx a = v[0];
            cplx b1 = v[1] + v[4], b2 = v[2] + v[3];
            cplx e1 = v[1] - v[4], e2 = v[2] - v[3];
            v[0] = a + b1 + b2;
            cplx m1 = a + c1 * b1 + c2 * b2;
            cplx m2 = a + c2 * b1 + c1 * b2;
            cplx n1 = -I * (s1 * e1 + s2 * e2);
            cplx n2 = -I * (s2 * e1 - s1 * e2);
            v[1] = m1 + n1; v[4] = m1 - n1;
            v[2] = m2 + n2; v[3] = m2 - n2;
        } break;
        default: {
            const size_t rstep = n / (size_t)R;
            for (int k = 0; k < R; ++k) {
                cplx acc = 0;
                for (int r = 0; r < R; ++r)
                    acc += v[r] * p->tw[(((size_t)k * (size_t)r) % (size_t)R) * rstep];
                t[k] = acc;
            }
            for (int k = 0; k < R; ++k) v[k] = t[k];
        } break;
        }
        const size_t base = (j / Ns) * Ns * (size_t)R + jm;
        for (int r = 0; r < R; ++r) out[base + (size_t)r * Ns] = v[r];
    }
}

/* forward transform in place; work has room for n (or m for Bluestein) values */
static void execute(const orc_fft_plan *p, cplx *data, cplx *work)
{
    const size_t n = p->n;
    if (n <= 1) return;
    if (p->bluestein) {
        const size_t m = p->m;
        cplx *a = (cplx *)calloc(m, sizeof(cplx));
        cplx *w = (cplx *)malloc(sizeof(cplx) * m);
        for (size_t k = 0; k < n; ++k) a[k] = data[k] * p->chirp[k];
        execute(p->sub, a, w);
        for (size_t k = 0; k < m; ++k) a[k] *= p->chirp_fft[k];
        for (size_t k = 0; k < m; ++k) a[k] = conj(a[k]);
        execute(p->sub, a, w);
        for (size_t k = 0; k < n; ++k) data[k] = conj(a[k]) / (double)m * p->chirp[k];
        free(a);
        free(w);
        return;
    }
    cplx *src = data, *dst = work;
    size_t Ns = 1;
    for (int f = 0; f < p->nf; ++f) {
        stockham_pass(p, p->fac[f], Ns, src, dst);
        Ns *= (size_t)p->fac[f];
        cplx *tmp = src; src = dst; dst = tmp;
    }
    if (src != data) memcpy(data, src, sizeof(cplx) * n);
}

int orc_fft(size_t n, const double *in, double *out, int sign)
{
    if (n == 0) return ORC_EINVAL;
    orc_fft_plan *p = plan_new(n);
    cplx *d = (cplx *)malloc(sizeof(cplx) * n);
    cplx *w = (cplx *)malloc(sizeof(cplx) * n);
    if (!p || !d || !w) { plan_free(p); free(d); free(w); return ORC_ENOMEM; }
    for (size_t k = 0; k < n; ++k) {
        d[k] = in[2 * k] + I * in[2 * k + 1];
        if (sign > 0) d[k] = conj(d[k]);
    }
    execute(p, d, w);
    for (size_t k = 0; k < n; ++k) {
        cplx z = sign > 0 ? conj(d[k]) : d[k];
        out[2 * k] = creal(z);
        out[2 * k + 1] = cimag(z);
    }
    plan_free(p); free(d); free(w);
    return ORC_OK;
}

/* reusable real-transform plan */
struct orc_rfft {
    size_t n;           /* real length (even) */
    orc_fft_plan *half; /* complex length n/2 */
    cplx *post;         /* e^{-2 pi i k / n}, k <= n/2 */
    cplx *buf, *work;
};

orc_rfft *orc_rfft_new(size_t n)
{
    if (n < 2 || (n & 1)) return NULL;
    orc_rfft *r = (orc_rfft *)calloc(1, sizeof(*r));
    if (!r) return NULL;
    r->n = n;
    const size_t h = n / 2;
    r->half = plan_new(h);
    r->post = (cplx *)malloc(sizeof(cplx) * (h + 1));
    size_t wl = r->half && r->half->bluestein ? r->half->m : h;
    r->buf = (cplx *)malloc(sizeof(cplx) * (h + 1));
    r->work = (cplx *)malloc(sizeof(cplx) * wl);
    if (!r->half || !r->post || !r->buf || !r->work) { orc_rfft_free(r); return NULL; }
    for (size_t k = 0; k <= h; ++k) r->post[k] = twiddle_ld(k, n);
    return r;
}

void orc_rfft_free(orc_rfft *r)
{
    if (!r) return;
    plan_free(r->half);
    free(r->post);
    free(r->buf);
    free(r->work);
    free(r);
}

/* X_k, k = 0..n/2, interleaved re/im */
void orc_rfft_forward(orc_rfft *r, const double *x, double *X)
{
    const size_t h = r->n / 2;
    cplx *z = r->buf;
    for (size_t j = 0; j < h; ++j) z[j] = x[2 * j] + I * x[2 * j + 1];
    execute(r->half, z, r->work);
    z[h] = z[0];
    for (size_t k = 0; k <= h / 2; ++k) {
        const size_t kk = h - k;
        cplx a = z[k], b = conj(z[kk]);
        cplx e = 0.5 * (a + b), o = -0.5 * I * (a - b);
        cplx Xk = e + r->post[k] * o;
        cplx a2 = z[kk], b2 = conj(z[k]);
        cplx e2 = 0.5 * (a2 + b2), o2 = -0.5 * I * (a2 - b2);
        cplx Xkk = e2 + r->post[kk] * o2;
        X[2 * k] = creal(Xk); X[2 * k + 1] = cimag(Xk);
        X[2 * kk] = creal(Xkk); X[2 * kk + 1] = cimag(Xkk);
    }
}

/* unnormalised Hermitian backward transform; Im X_0 and Im X_{n/2} are ignored */
void orc_rfft_backward(orc_rfft *r, const double *X, double *x)
{
    const size_t h = r->n / 2;
    cplx *z = r->buf;
    for (size_t k = 0; k < h; ++k) {
        const size_t kk = h - k;
        cplx a = X[2 * k] + I * ((k == 0) ? 0.0 : X[2 * k + 1]);
        cplx b = X[2 * kk] - I * ((kk == h) ? 0.0 : X[2 * kk + 1]);   /* conj(X_{h-k}) */
        cplx e = a + b;
        cplx o = (a - b) * conj(r->post[k]);
        z[k] = conj(e + I * o);    /* conj trick: backward = conj(forward(conj)) */
    }
    execute(r->half, z, r->work);
    for (size_t j = 0; j < h; ++j) {
        cplx v = conj(z[j]);
        x[2 * j] = creal(v);
        x[2 * j + 1] = cimag(v);
    }
}

int orc_rfft_r2c(size_t n, const double *x, double *X)
{
    orc_rfft *r = orc_rfft_new(n);
    if (!r) return ORC_EINVAL;
    orc_rfft_forward(r, x, X);
    orc_rfft_free(r);
    return ORC_OK;
}

int orc_rfft_c2r(size_t n, const double *X, double *x)
{
    orc_rfft *r = orc_rfft_new(n);
    if (!r) return ORC_EINVAL;
    orc_rfft_backward(r, X, x);
    orc_rfft_free(r);
    return ORC_OK;
}

/* 2D row-major n1 x n2 real -> n1 x (n2/2+1) complex, FFTW r2c_2d semantics.  Rows,
   then columns, each transform independent: OpenMP threads take rows and columns with
   plans and buffers of their own (the arithmetic per transform is the serial one) */
int orc_rfft2_r2c(size_t n1, size_t n2, const double *x, double *X)
{
    const size_t h2 = n2 / 2 + 1;
    int fail = 0;
#pragma omp parallel
    {
        orc_rfft *r = orc_rfft_new(n2);
        orc_fft_plan *c = plan_new(n1);
        cplx *col = (cplx *)malloc(sizeof(cplx) * n1);
        cplx *w = (cplx *)malloc(sizeof(cplx) * ((c && c->bluestein) ? c->m : n1));
        const int ok = r && c && col && w;
        if (!ok) {
#pragma omp atomic write
            fail = 1;
        }
#pragma omp for schedule(static)
        for (long i = 0; i < (long)n1; ++i)
            if (ok) orc_rfft_forward(r, x + (size_t)i * n2, X + 2 * (size_t)i * h2);
#pragma omp for schedule(static)
        for (long jj = 0; jj < (long)h2; ++jj) {
            if (!ok) continue;
            const size_t j = (size_t)jj;
            for (size_t i = 0; i < n1; ++i) col[i] = X[2 * (i * h2 + j)] + I * X[2 * (i * h2 + j) + 1];
            execute(c, col, w);
            for (size_t i = 0; i < n1; ++i) { X[2 * (i * h2 + j)] = creal(col[i]); X[2 * (i * h2 + j) + 1] = cimag(col[i]); }
        }
        orc_rfft_free(r); plan_free(c); free(col); free(w);
    }
    return fail ? ORC_ENOMEM : ORC_OK;
}

/* inverse of the above (unnormalised, FFTW c2r_2d semantics: the last dimension is
   Hermitian-completed, imaginary parts of its DC and Nyquist columns are dropped
   after the column transforms); threads as above */
int orc_rfft2_c2r(size_t n1, size_t n2, const double *X, double *x)
{
    const size_t h2 = n2 / 2 + 1;
    double *tmp = (double *)malloc(sizeof(double) * 2 * n1 * h2);
    if (!tmp) return ORC_ENOMEM;
    int fail = 0;
#pragma omp parallel
    {
        orc_rfft *r = orc_rfft_new(n2);
        orc_fft_plan *c = plan_new(n1);
        cplx *col = (cplx *)malloc(sizeof(cplx) * n1);
        cplx *w = (cplx *)malloc(sizeof(cplx) * ((c && c->bluestein) ? c->m : n1));
        const int ok = r && c && col && w;
        if (!ok) {
#pragma omp atomic write
            fail = 1;
        }
#pragma omp for schedule(static)
        for (long jj = 0; jj < (long)h2; ++jj) {
            if (!ok) continue;
            const size_t j = (size_t)jj;
            for (size_t i = 0; i < n1; ++i) col[i] = X[2 * (i * h2 + j)] - I * X[2 * (i * h2 + j) + 1];
            execute(c, col, w);
            for (size_t i = 0; i < n1; ++i) { tmp[2 * (i * h2 + j)] = creal(col[i]); tmp[2 * (i * h2 + j) + 1] = -cimag(col[i]); }
        }
#pragma omp for schedule(static)
        for (long i = 0; i < (long)n1; ++i)
            if (ok) orc_rfft_backward(r, tmp + 2 * (size_t)i * h2, x + (size_t)i * n2);
        orc_rfft_free(r); plan_free(c); free(col); free(w);
    }
    free(tmp);
    return fail ? ORC_ENOMEM : ORC_OK;
}
