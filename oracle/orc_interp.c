/*
 * Oracle (test infrastructure only): 1D and 2D interpolation restated from the
 * published methods that GSL 2.7 implements and COFFE calls.
 *
 *  - natural cubic spline (twofast.c:399-401 resamples P(k) with it), solved with
 *    the symmetric tridiagonal LDL^T recursion; c_i = y''(x_i)/2, evaluation
 *    y = y_i + dx (b_i + dx (c_i + dx d_i)), b_i = dy/h - h (c_{i+1} + 2 c_i)/3,
 *    d_i = (c_{i+1} - c_i)/(3 h);
 *  - Akima 1970 without rounding at the ends (the default `interp_method`,
 *    parser.c:807; used through common.c:524-576 and common.c:704-710): slopes
 *    m_i of the data, end slopes extended linearly (m_{-1} = 2 m_0 - m_1,
 *    m_{-2} = 3 m_0 - 2 m_1 and mirrored at the top), node derivative
 *    t_i = ((|m_{i+1}-m_i|) m_{i-1} + |m_{i-1}-m_{i-2}| m_i)/(sum), or m_{i-1}...
 *    written as the weights alpha used below, with a flat node when both
 *    weights vanish;
 *  - binary search with the x_lo <= x < x_hi convention (top node maps to the last
 *    interval); out-of-range evaluations return NaN (GSL's EDOM value);
 *  - bicubic Hermite patches whose node derivatives come from natural cubic
 *    splines along x, along y and (cross term) along y of the x-derivatives
 *    (covariance.c:443-464 through common.c:579-613, 712-719).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

size_t orc_bsearch(const double *xa, double x, size_t lo, size_t hi)
{
    while (hi > lo + 1) {
        size_t i = (hi + lo) / 2;
        if (xa[i] > x) hi = i;
        else lo = i;
    }
    return lo;
}

/* symmetric tridiagonal solve: diag[n], off[n-1], rhs[n] -> x[n] (LDL^T) */
static int symm_tridiag(const double *diag, const double *off, const double *rhs, double *x, size_t n)
{
    double *gamma = (double *)malloc(sizeof(double) * n);
    double *alpha = (double *)malloc(sizeof(double) * n);
    double *z = (double *)malloc(sizeof(double) * n);
    if (!gamma || !alpha || !z) { free(gamma); free(alpha); free(z); return ORC_ENOMEM; }
    alpha[0] = diag[0];
    gamma[0] = n > 1 ? off[0] / alpha[0] : 0.0;
    for (size_t i = 1; i + 1 < n; ++i) {
        alpha[i] = diag[i] - off[i - 1] * gamma[i - 1];
        gamma[i] = off[i] / alpha[i];
    }
    if (n > 1) alpha[n - 1] = diag[n - 1] - off[n - 2] * gamma[n - 2];
    z[0] = rhs[0];
    for (size_t i = 1; i < n; ++i) z[i] = rhs[i] - gamma[i - 1] * z[i - 1];
    for (size_t i = 0; i < n; ++i) z[i] = z[i] / alpha[i];
    x[n - 1] = z[n - 1];
    for (size_t i = n - 1; i-- > 0;) x[i] = z[i] - gamma[i] * x[i + 1];
    free(gamma); free(alpha); free(z);
    return ORC_OK;
}

int orc_cspline_init(size_t n, const double *x, const double *y, double *c)
{
    if (n < 3) return ORC_EINVAL;
    const size_t sys = n - 2;
    c[0] = 0.0;
    c[n - 1] = 0.0;
    double *diag = (double *)malloc(sizeof(double) * sys);
    double *off = (double *)malloc(sizeof(double) * sys);
    double *g = (double *)malloc(sizeof(double) * sys);
    if (!diag || !off || !g) { free(diag); free(off); free(g); return ORC_ENOMEM; }
    for (size_t i = 0; i < sys; ++i) {
        const double h_i = x[i + 1] - x[i];
        const double h_ip1 = x[i + 2] - x[i + 1];
        const double g_i = h_i != 0.0 ? 1.0 / h_i : 0.0;
        const double g_ip1 = h_ip1 != 0.0 ? 1.0 / h_ip1 : 0.0;
        off[i] = h_ip1;
        diag[i] = 2.0 * (h_ip1 + h_i);
        g[i] = 3.0 * ((y[i + 2] - y[i + 1]) * g_ip1 - (y[i + 1] - y[i]) * g_i);
    }
    int st;
    if (sys == 1) { c[1] = g[0] / diag[0]; st = ORC_OK; }
    else st = symm_tridiag(diag, off, g, c + 1, sys);
    free(diag); free(off); free(g);
    return st;
}

double orc_cspline_eval(size_t n, const double *x, const double *y, const double *c, double xq)
{
    if (!(xq >= x[0] && xq <= x[n - 1])) return NAN;
    const size_t i = orc_bsearch(x, xq, 0, n - 1);
    const double dx = x[i + 1] - x[i];
    const double dy = y[i + 1] - y[i];
    const double delx = xq - x[i];
    const double b = dy / dx - dx * (c[i + 1] + 2.0 * c[i]) / 3.0;
    const double d = (c[i + 1] - c[i]) / (3.0 * dx);
    return y[i] + delx * (b + delx * (c[i] + delx * d));
}

double orc_cspline_deriv(size_t n, const double *x, const double *y, const double *c, double xq)
{
    if (!(xq >= x[0] && xq <= x[n - 1])) return NAN;
    const size_t i = orc_bsearch(x, xq, 0, n - 1);
    const double dx = x[i + 1] - x[i];
    const double dy = y[i + 1] - y[i];
    const double delx = xq - x[i];
    const double b = dy / dx - dx * (c[i + 1] + 2.0 * c[i]) / 3.0;
    const double d = (c[i + 1] - c[i]) / (3.0 * dx);
    return b + delx * (2.0 * c[i] + 3.0 * d * delx);
}

int orc_akima_init(size_t n, const double *x, const double *y, double *b, double *c, double *d)
{
    if (n < 5) return ORC_EINVAL;
    double *mm = (double *)malloc(sizeof(double) * (n + 4));
    if (!mm) return ORC_ENOMEM;
    double *m = mm + 2;   /* m[-2] .. m[n+1] */
    for (size_t i = 0; i + 1 < n; ++i) m[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
    m[-2] = 3.0 * m[0] - 2.0 * m[1];
    m[-1] = 2.0 * m[0] - m[1];
    m[n - 1] = 2.0 * m[n - 2] - m[n - 3];
    m[n] = 3.0 * m[n - 2] - 2.0 * m[n - 3];
    for (size_t ii = 0; ii + 1 < n; ++ii) {
        const long i = (long)ii;
        const double ne = fabs(m[i + 1] - m[i]) + fabs(m[i - 1] - m[i - 2]);
        if (ne == 0.0) {
            b[i] = m[i];
            c[i] = 0.0;
            d[i] = 0.0;
        } else {
            const double h = x[i + 1] - x[i];
            const double ne_next = fabs(m[i + 2] - m[i + 1]) + fabs(m[i] - m[i - 1]);
            const double alpha = fabs(m[i - 1] - m[i - 2]) / ne;
            double t_next;
            if (ne_next == 0.0) {
                t_next = m[i];
            } else {
                const double alpha_next = fabs(m[i] - m[i - 1]) / ne_next;
                t_next = (1.0 - alpha_next) * m[i] + alpha_next * m[i + 1];
            }
            b[i] = (1.0 - alpha) * m[i - 1] + alpha * m[i];
            c[i] = (3.0 * m[i] - 2.0 * b[i] - t_next) / h;
            d[i] = (b[i] + t_next - 2.0 * m[i]) / (h * h);
        }
    }
    free(mm);
    return ORC_OK;
}

double orc_akima_eval(size_t n, const double *x, const double *y, const double *b, const double *c,
                      const double *d, double xq)
{
    if (!(xq >= x[0] && xq <= x[n - 1])) return NAN;
    const size_t i = orc_bsearch(x, xq, 0, n - 1);
    const double delx = xq - x[i];
    return y[i] + delx * (b[i] + delx * (c[i] + d[i] * delx));
}

/* ---- forward-mode derivative of the Akima interpolant in its node values ------ */
typedef struct { double v, d; } orc_dual;

static orc_dual dd_sub(orc_dual a, orc_dual b) { orc_dual r = {a.v - b.v, a.d - b.d}; return r; }
static orc_dual dd_add(orc_dual a, orc_dual b) { orc_dual r = {a.v + b.v, a.d + b.d}; return r; }
static orc_dual dd_scale(double s, orc_dual a) { orc_dual r = {s * a.v, s * a.d}; return r; }
static orc_dual dd_mul(orc_dual a, orc_dual b) { orc_dual r = {a.v * b.v, a.d * b.v + a.v * b.d}; return r; }
static orc_dual dd_div(orc_dual a, orc_dual b)
{
    orc_dual r = {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)};
    return r;
}
static orc_dual dd_abs(orc_dual a)
{
    orc_dual r = {fabs(a.v), a.v > 0 ? a.d : (a.v < 0 ? -a.d : 0.0)};
    return r;
}

/*
 * The Akima interpolant of orc_akima_init / orc_akima_eval at xq[q] (out) and its
 * directional derivative with respect to the node values y along dy (dout): every
 * step of the slope formula carried as a (value, derivative) pair, |u| with
 * derivative sign(u) du (the interpolant is differentiable wherever no weight
 * |m_{i+1} - m_i| vanishes).
 */
int orc_akima_jvp(size_t n, const double *x, const double *y, const double *dy, size_t nq, const double *xq,
                  double *out, double *dout)
{
    if (n < 5) return ORC_EINVAL;
    orc_dual *mm = (orc_dual *)malloc(sizeof(orc_dual) * (n + 4));
    if (!mm) return ORC_ENOMEM;
    orc_dual *m = mm + 2;
    for (size_t i = 0; i + 1 < n; ++i) {
        const double h = x[i + 1] - x[i];
        m[i].v = (y[i + 1] - y[i]) / h;
        m[i].d = (dy[i + 1] - dy[i]) / h;
    }
    m[-2] = dd_sub(dd_scale(3.0, m[0]), dd_scale(2.0, m[1]));
    m[-1] = dd_sub(dd_scale(2.0, m[0]), m[1]);
    m[n - 1] = dd_sub(dd_scale(2.0, m[n - 2]), m[n - 3]);
    m[n] = dd_sub(dd_scale(3.0, m[n - 2]), dd_scale(2.0, m[n - 3]));
    for (size_t q = 0; q < nq; ++q) {
        if (!(xq[q] >= x[0] && xq[q] <= x[n - 1])) { out[q] = NAN; dout[q] = NAN; continue; }
        const long i = (long)orc_bsearch(x, xq[q], 0, n - 1);
        const double h = x[i + 1] - x[i];
        const orc_dual ne = dd_add(dd_abs(dd_sub(m[i + 1], m[i])), dd_abs(dd_sub(m[i - 1], m[i - 2])));
        orc_dual b, c, d;
        if (ne.v == 0.0) {
            b = m[i];
            c.v = c.d = d.v = d.d = 0.0;
        } else {
            const orc_dual ne_next = dd_add(dd_abs(dd_sub(m[i + 2], m[i + 1])), dd_abs(dd_sub(m[i], m[i - 1])));
            const orc_dual alpha = dd_div(dd_abs(dd_sub(m[i - 1], m[i - 2])), ne);
            const orc_dual one = {1.0, 0.0};
            orc_dual t_next;
            if (ne_next.v == 0.0) {
                t_next = m[i];
            } else {
                const orc_dual an = dd_div(dd_abs(dd_sub(m[i], m[i - 1])), ne_next);
                t_next = dd_add(dd_mul(dd_sub(one, an), m[i]), dd_mul(an, m[i + 1]));
            }
            b = dd_add(dd_mul(dd_sub(one, alpha), m[i - 1]), dd_mul(alpha, m[i]));
            c = dd_scale(1.0 / h, dd_sub(dd_sub(dd_scale(3.0, m[i]), dd_scale(2.0, b)), t_next));
            d = dd_scale(1.0 / (h * h), dd_sub(dd_add(b, t_next), dd_scale(2.0, m[i])));
        }
        const double delx = xq[q] - x[i];
        out[q] = y[i] + delx * (b.v + delx * (c.v + d.v * delx));
        dout[q] = dy[i] + delx * (b.d + delx * (c.d + d.d * delx));
    }
    free(mm);
    return ORC_OK;
}

/* convenience: build + evaluate at many points */
int orc_akima_interp(size_t n, const double *x, const double *y, size_t nq, const double *xq, double *out)
{
    double *b = (double *)malloc(sizeof(double) * n * 3);
    if (!b) return ORC_ENOMEM;
    int st = orc_akima_init(n, x, y, b, b + n, b + 2 * n);
    if (st == ORC_OK)
        for (size_t i = 0; i < nq; ++i) out[i] = orc_akima_eval(n, x, y, b, b + n, b + 2 * n, xq[i]);
    free(b);
    return st;
}

int orc_cspline_interp(size_t n, const double *x, const double *y, size_t nq, const double *xq, double *out)
{
    double *c = (double *)malloc(sizeof(double) * n);
    if (!c) return ORC_ENOMEM;
    int st = orc_cspline_init(n, x, y, c);
    if (st == ORC_OK)
        for (size_t i = 0; i < nq; ++i) out[i] = orc_cspline_eval(n, x, y, c, xq[i]);
    free(c);
    return st;
}

/* ---- bicubic ------------------------------------------------------------- */
struct orc_bicubic {
    size_t nx, ny;
    const double *x, *y, *z;     /* z[j*nx + i] = f(x_i, y_j) (borrowed) */
    double *zx, *zy, *zxy;       /* same layout */
};

static int spline_derivs(size_t n, const double *x, const double *v, size_t stride, double *out, size_t ostride,
                         double *ybuf, double *cbuf)
{
    for (size_t i = 0; i < n; ++i) ybuf[i] = v[i * stride];
    int st = orc_cspline_init(n, x, ybuf, cbuf);
    if (st != ORC_OK) return st;
    for (size_t i = 0; i < n; ++i) out[i * ostride] = orc_cspline_deriv(n, x, ybuf, cbuf, x[i]);
    return ORC_OK;
}

orc_bicubic *orc_bicubic_new(size_t nx, size_t ny, const double *x, const double *y, const double *z)
{
    orc_bicubic *s = (orc_bicubic *)calloc(1, sizeof(*s));
    if (!s) return NULL;
    s->nx = nx; s->ny = ny; s->x = x; s->y = y; s->z = z;
    s->zx = (double *)malloc(sizeof(double) * nx * ny);
    s->zy = (double *)malloc(sizeof(double) * nx * ny);
    s->zxy = (double *)malloc(sizeof(double) * nx * ny);
    if (!s->zx || !s->zy || !s->zxy) { orc_bicubic_free(s); return NULL; }
    const size_t mx = nx > ny ? nx : ny;
    int fail = 0;
    /* every row / column spline is independent: OpenMP threads with buffers of their own */
#pragma omp parallel
    {
        double *yb = (double *)malloc(sizeof(double) * mx);
        double *cb = (double *)malloc(sizeof(double) * mx);
        const int ok = yb && cb;
        if (!ok) {
#pragma omp atomic write
            fail = 1;
        }
#pragma omp for schedule(static)
        for (long j = 0; j < (long)ny; ++j)
            if (ok) spline_derivs(nx, x, z + (size_t)j * nx, 1, s->zx + (size_t)j * nx, 1, yb, cb);
#pragma omp for schedule(static)
        for (long i = 0; i < (long)nx; ++i)
            if (ok) spline_derivs(ny, y, z + i, nx, s->zy + i, nx, yb, cb);
#pragma omp for schedule(static)
        for (long i = 0; i < (long)nx; ++i)
            if (ok) spline_derivs(ny, y, s->zx + i, nx, s->zxy + i, nx, yb, cb);
        free(yb); free(cb);
    }
    if (fail) { orc_bicubic_free(s); return NULL; }
    return s;
}

void orc_bicubic_free(orc_bicubic *s)
{
    if (!s) return;
    free(s->zx); free(s->zy); free(s->zxy);
    free(s);
}

double orc_bicubic_eval(const orc_bicubic *s, double xq, double yq)
{
    const size_t nx = s->nx, ny = s->ny;
    if (!(xq >= s->x[0] && xq <= s->x[nx - 1] && yq >= s->y[0] && yq <= s->y[ny - 1])) return NAN;
    const size_t i = orc_bsearch(s->x, xq, 0, nx - 1);
    const size_t j = orc_bsearch(s->y, yq, 0, ny - 1);
    const double dx = s->x[i + 1] - s->x[i], dy = s->y[j + 1] - s->y[j];
    const double t = (xq - s->x[i]) / dx, u = (yq - s->y[j]) / dy;
#define Z(A, a, b) (A)[((j) + (b)) * nx + (i) + (a)]
    /* corner values and unit-cell derivatives */
    double f[2][2], fx[2][2], fy[2][2], fxy[2][2];
    for (int a = 0; a < 2; ++a)
        for (int bb = 0; bb < 2; ++bb) {
            f[a][bb] = Z(s->z, a, bb);
            fx[a][bb] = Z(s->zx, a, bb) * dx;
            fy[a][bb] = Z(s->zy, a, bb) * dy;
            fxy[a][bb] = Z(s->zxy, a, bb) * dx * dy;
        }
#undef Z
    /* cubic Hermite basis */
    const double ht[4] = {
        (1 + 2 * t) * (1 - t) * (1 - t),   /* h00 */
        t * t * (3 - 2 * t),               /* h01 */
        t * (1 - t) * (1 - t),             /* h10 */
        t * t * (t - 1)                    /* h11 */
    };
    const double hu[4] = {
        (1 + 2 * u) * (1 - u) * (1 - u),
        u * u * (3 - 2 * u),
        u * (1 - u) * (1 - u),
        u * u * (u - 1)
    };
    double r = 0.0;
    for (int a = 0; a < 2; ++a)
        for (int bb = 0; bb < 2; ++bb) {
            r += f[a][bb] * ht[a] * hu[bb];
            r += fx[a][bb] * ht[2 + a] * hu[bb];
            r += fy[a][bb] * ht[a] * hu[2 + bb];
            r += fxy[a][bb] * ht[2 + a] * hu[2 + bb];
        }
    return r;
}

/* the bicubic interpolant of z (z[j nx + i] = f(x_i, y_j)) at nq points (NaN outside) */
int orc_bicubic_interp(size_t nx, size_t ny, const double *x, const double *y, const double *z, size_t nq,
                       const double *xq, const double *yq, double *out)
{
    orc_bicubic *s = orc_bicubic_new(nx, ny, x, y, z);
    if (!s) return ORC_ENOMEM;
    for (size_t q = 0; q < nq; ++q) out[q] = orc_bicubic_eval(s, xq[q], yq[q]);
    orc_bicubic_free(s);
    return ORC_OK;
}
