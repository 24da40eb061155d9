/*
 * Host interpolation used by the per-call (drop-in) path and the Python API's
 * scalar lookups: Akima (1970, non-rounded ends), the default `interp_method`
 * (parser.c:807) through coffe_init_spline / coffe_interp_spline
 * (common.c:524-576, 704-710).  Out-of-range points give NaN.
 */
#include <math.h>
#include <stdlib.h>

#include "coffe_engine.h"

static size_t bsearch_lo(const double *x, double v, size_t lo, size_t hi)
{
    while (hi > lo + 1) {
        const size_t m = (hi + lo) / 2;
        if (x[m] > v) hi = m;
        else lo = m;
    }
    return lo;
}

int coffe_akima_eval(size_t n, const double *x, const double *y, size_t nq, const double *xq, double *out)
{
    if (n < 5 || !x || !y || (nq && (!xq || !out))) return COFFE_EINVAL;
    double *buf = (double *)malloc(sizeof(double) * (4 * n + 4));
    if (!buf) return COFFE_ENOMEM;
    double *m = buf + 2, *b = buf + n + 4, *c = b + n, *d = c + n;
    for (size_t i = 0; i + 1 < n; ++i) m[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
    m[-2] = 3.0 * m[0] - 2.0 * m[1];
    m[-1] = 2.0 * m[0] - m[1];
    m[n - 1] = 2.0 * m[n - 2] - m[n - 3];
    m[n] = 3.0 * m[n - 2] - 2.0 * m[n - 3];
    for (long i = 0; i + 1 < (long)n; ++i) {
        const double ne = fabs(m[i + 1] - m[i]) + fabs(m[i - 1] - m[i - 2]);
        if (ne == 0.0) {
            b[i] = m[i]; c[i] = 0.0; d[i] = 0.0;
            continue;
        }
        const double h = x[i + 1] - x[i];
        const double nn = fabs(m[i + 2] - m[i + 1]) + fabs(m[i] - m[i - 1]);
        const double al = fabs(m[i - 1] - m[i - 2]) / ne;
        const double tn = nn == 0.0 ? m[i] : (1.0 - fabs(m[i] - m[i - 1]) / nn) * m[i] + fabs(m[i] - m[i - 1]) / nn * m[i + 1];
        b[i] = (1.0 - al) * m[i - 1] + al * m[i];
        c[i] = (3.0 * m[i] - 2.0 * b[i] - tn) / h;
        d[i] = (b[i] + tn - 2.0 * m[i]) / (h * h);
    }
    for (size_t q = 0; q < nq; ++q) {
        const double v = xq[q];
        if (!(v >= x[0] && v <= x[n - 1])) { out[q] = NAN; continue; }
        const size_t i = bsearch_lo(x, v, 0, n - 1);
        const double dx = v - x[i];
        out[q] = y[i] + dx * (b[i] + dx * (c[i] + d[i] * dx));
    }
    free(buf);
    return COFFE_OK;
}
