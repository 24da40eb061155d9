/*
 * Device helpers shared by the HIP sources (gfx950, FP64).
 */
#ifndef COFFE_ENGINE_DEVICE_H
#define COFFE_ENGINE_DEVICE_H

#include <hip/hip_runtime.h>

namespace ce {

// interval j of the sorted table x[0..n) with x[j] <= v < x[j+1] (j = n - 2 at
// v = x[n-1]), the result of a bisection over [0, n-1]; v must lie in [x[0], x[n-1]]
__device__ inline int interval_bisect(const double *__restrict__ x, int lo, int hi, double v)
{
    while (hi > lo + 1) {
        const int m = (lo + hi) >> 1;
        if (x[m] > v) hi = m;
        else lo = m;
    }
    return lo;
}

// the same interval, starting from a guess g: a few steps of walking settle a good
// guess (exact for a uniform grid and a linear guess); a poor one falls back to
// bisection on the bracket the walk has established
__device__ inline int interval_from(const double *__restrict__ x, int n, double v, int g)
{
    g = g < 0 ? 0 : (g > n - 2 ? n - 2 : g);
#pragma unroll 1
    for (int it = 0; it < 6; ++it) {
        if (x[g] > v) {
            if (g == 0) return 0;
            --g;
        } else if (g < n - 2 && x[g + 1] <= v) {
            ++g;
        } else {
            return g;
        }
    }
    return (x[g] > v) ? interval_bisect(x, 0, g, v) : interval_bisect(x, g, n - 1, v);
}

// a real-valued index estimate as an int in [0, n - 1] (NaN and infinities included)
__device__ inline int clamp_index(double t, int n)
{
    t = fmin(fmax(t, 0.0), (double)(n - 1));
    return (t == t) ? (int)t : 0;
}

// linear guess of the interval of v on x[0..n) (exact for a uniform grid)
__device__ inline int interval_guess(const double *__restrict__ x, int n, double v)
{
    return clamp_index((v - x[0]) / (x[n - 1] - x[0]) * (double)(n - 1), n);
}

// Akima value at xq in interval j of the table (x, y) of length n (n >= 5), GSL
// non-rounded ends (slopes beyond the ends extended linearly)
__device__ inline double akima_at(const double *__restrict__ x, const double *__restrict__ y, int n, int j, double xq)
{
    auto slope = [&](int a) -> double {
        if (a >= 0 && a <= n - 2) return (y[a + 1] - y[a]) / (x[a + 1] - x[a]);
        const double m0 = (y[1] - y[0]) / (x[1] - x[0]);
        const double m1 = (y[2] - y[1]) / (x[2] - x[1]);
        const double mn2 = (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]);
        const double mn3 = (y[n - 2] - y[n - 3]) / (x[n - 2] - x[n - 3]);
        if (a == -1) return 2.0 * m0 - m1;
        if (a == -2) return 3.0 * m0 - 2.0 * m1;
        if (a == n - 1) return 2.0 * mn2 - mn3;
        return 3.0 * mn2 - 2.0 * mn3;   // a == n
    };
    const double mm2 = slope(j - 2), mm1 = slope(j - 1), m0 = slope(j), mp1 = slope(j + 1), mp2 = slope(j + 2);
    const double ne = fabs(mp1 - m0) + fabs(mm1 - mm2);
    const double delx = xq - x[j];
    if (ne == 0.0) return y[j] + delx * m0;
    const double h = x[j + 1] - x[j];
    const double nen = fabs(mp2 - mp1) + fabs(m0 - mm1);
    const double alpha = fabs(mm1 - mm2) / ne;
    double tn;
    if (nen == 0.0) {
        tn = m0;
    } else {
        const double an = fabs(m0 - mm1) / nen;
        tn = (1.0 - an) * m0 + an * mp1;
    }
    const double b = (1.0 - alpha) * mm1 + alpha * m0;
    const double c = (3.0 * m0 - 2.0 * b - tn) / h;
    const double d = (b + tn - 2.0 * m0) / (h * h);
    return y[j] + delx * (b + delx * (c + d * delx));
}

// Akima value at xq on the table (x, y) of length n (n >= 5); NaN outside [x[0], x[n-1]]
__device__ inline double akima_general(const double *__restrict__ x, const double *__restrict__ y, int n, double xq)
{
    if (!(xq >= x[0] && xq <= x[n - 1])) return __builtin_nan("");
    return akima_at(x, y, n, interval_bisect(x, 0, n - 1, xq), xq);
}

}  // namespace ce

#endif
