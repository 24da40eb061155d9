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

// forward-mode dual number: a value and its derivative along one direction (the
// tangent), for the JVP instances of the kernels; arithmetic on the value part is the
// same as on doubles, so a dual evaluation reproduces the double one in its value
struct dd {
    double v, t;
    dd() = default;
    __device__ dd(double v_) : v(v_), t(0.0) {}
    __device__ dd(double v_, double t_) : v(v_), t(t_) {}
};
__device__ inline dd operator+(dd a, dd b) { return dd(a.v + b.v, a.t + b.t); }
__device__ inline dd operator-(dd a, dd b) { return dd(a.v - b.v, a.t - b.t); }
__device__ inline dd operator-(dd a) { return dd(-a.v, -a.t); }
__device__ inline dd operator*(dd a, dd b) { return dd(a.v * b.v, a.t * b.v + a.v * b.t); }
__device__ inline dd operator/(dd a, dd b)
{
    const double q = a.v / b.v;
    return dd(q, (a.t - q * b.t) / b.v);
}
__device__ inline dd operator+(dd a, double b) { return dd(a.v + b, a.t); }
__device__ inline dd operator+(double a, dd b) { return dd(a + b.v, b.t); }
__device__ inline dd operator-(dd a, double b) { return dd(a.v - b, a.t); }
__device__ inline dd operator-(double a, dd b) { return dd(a - b.v, -b.t); }
__device__ inline dd operator*(dd a, double b) { return dd(a.v * b, a.t * b); }
__device__ inline dd operator*(double a, dd b) { return dd(a * b.v, a * b.t); }
__device__ inline dd operator/(dd a, double b) { return dd(a.v / b, a.t / b); }
__device__ inline dd operator/(double a, dd b)
{
    const double q = a / b.v;
    return dd(q, -q * b.t / b.v);
}
__device__ inline dd &operator+=(dd &a, dd b) { a = a + b; return a; }
__device__ inline dd &operator-=(dd &a, dd b) { a = a - b; return a; }
__device__ inline dd &operator*=(dd &a, dd b) { a = a * b; return a; }
// |a| for both scalar types (for a dual, the derivative's sign taken at a = 0 from the
// + side)
__device__ inline double absv(double a) { return ::fabs(a); }
__device__ inline dd absv(dd a) { return a.v < 0.0 ? -a : a; }
__device__ inline double val(double a) { return a; }
__device__ inline double val(dd a) { return a.v; }
template <typename T> __device__ inline T make_dual(double v, double t);
template <> __device__ inline double make_dual<double>(double v, double) { return v; }
template <> __device__ inline dd make_dual<dd>(double v, double t) { return dd(v, t); }

// Akima (GSL, non-rounded ends) coefficients y_j, b, c, d of interval j of the table
// (x, y), optionally with the tangents xt, yt of the nodes (T = dd; null: zero): the
// value of the interval is y_j + delx (b + delx (c + d delx)), delx = xq - x_j
template <typename T>
struct akima_cf {
    T y0, b, c, d;
};
template <typename T>
__device__ inline akima_cf<T> akima_coefs(const double *__restrict__ x, const double *__restrict__ xt,
                                          const double *__restrict__ y, const double *__restrict__ yt, int n, int j)
{
    auto X = [&](int i) -> T { return make_dual<T>(x[i], xt ? xt[i] : 0.0); };
    auto Y = [&](int i) -> T { return make_dual<T>(y[i], yt ? yt[i] : 0.0); };
    auto slope = [&](int i) -> T {
        if (i >= 0 && i <= n - 2) return (Y(i + 1) - Y(i)) / (X(i + 1) - X(i));
        const T m0 = (Y(1) - Y(0)) / (X(1) - X(0));
        const T m1 = (Y(2) - Y(1)) / (X(2) - X(1));
        const T mn2 = (Y(n - 1) - Y(n - 2)) / (X(n - 1) - X(n - 2));
        const T mn3 = (Y(n - 2) - Y(n - 3)) / (X(n - 2) - X(n - 3));
        if (i == -1) return 2.0 * m0 - m1;
        if (i == -2) return 3.0 * m0 - 2.0 * m1;
        if (i == n - 1) return 2.0 * mn2 - mn3;
        return 3.0 * mn2 - 2.0 * mn3;
    };
    const T mm2 = slope(j - 2), mm1 = slope(j - 1), m0 = slope(j), mp1 = slope(j + 1), mp2 = slope(j + 2);
    const T ne = absv(mp1 - m0) + absv(mm1 - mm2);
    akima_cf<T> r;
    r.y0 = Y(j);
    if (val(ne) == 0.0) {
        r.b = m0;
        r.c = 0.0;
        r.d = 0.0;
    } else {
        const T h = X(j + 1) - X(j);
        const T nen = absv(mp2 - mp1) + absv(m0 - mm1);
        const T alpha = absv(mm1 - mm2) / ne;
        T tn;
        if (val(nen) == 0.0) {
            tn = m0;
        } else {
            const T an = absv(m0 - mm1) / nen;
            tn = (1.0 - an) * m0 + an * mp1;
        }
        r.b = (1.0 - alpha) * mm1 + alpha * m0;
        r.c = (3.0 * m0 - 2.0 * r.b - tn) / h;
        r.d = (r.b + tn - 2.0 * m0) / (h * h);
    }
    return r;
}

// Akima value at xq on the table (x, y) of length n (n >= 5); NaN outside [x[0], x[n-1]]
__device__ inline double akima_general(const double *__restrict__ x, const double *__restrict__ y, int n, double xq)
{
    if (!(xq >= x[0] && xq <= x[n - 1])) return __builtin_nan("");
    return akima_at(x, y, n, interval_bisect(x, 0, n - 1, xq), xq);
}

}  // namespace ce

#endif
