// Small-separation nodes of the I_l^n tables (gfx950, FP64).
//
// Below the FFTLog grid the reference fills each table with 43 direct integrals
// (integrals.c:520-557, min_sep of integrals.c:111-126):
//   node 0:      r = 0, (1/2pi^2) int dk k^(2+l-n) P(k) / (2l+1)!!   if n >= l, else 0
//                (integrals.c:238-269, coefficient 215-231)
//   node i > 0:  r = min_sep[i-1], (1/2pi^2) int dk k^2 P(k) j_l(kr)/(kr)^n, times r^(n-l)
//                if n >= l (integrals.c:280-325)
// over [kmin, kmax], P(k) the Akima spline of the input samples, by QUADPACK QAG
// with the 61-point Gauss-Kronrod rule, epsabs 0, epsrel 1e-5 (common.c:1004-1040).
//
// One wavefront runs one adaptive integration: lane j evaluates Kronrod node j of
// the interval being refined, the Gauss / Kronrod / |f| / |f - mean| sums are
// reduced with lane shuffles, and the interval list (a, b, result, error) lives in
// LDS.  Every lane holds the same reduced values, so the control flow of the
// bisection loop is uniform across the wavefront.
#include <float.h>
#include <hip/hip_runtime.h>

#include "ce_device.h"
#include "ce_internal.h"
#include "gk61_nodes.h"

namespace {

__constant__ double c_min_sep[CE_MIN_SEP_LEN];
__constant__ double c_xgk[31], c_wgk[31], c_wg[15];

// spherical Bessel j_l(x), integer l >= 0: power series for |x| < 0.5, upward
// recurrence from j_0, j_1 when x > l, otherwise Miller's downward recurrence
// normalised with sum_k (2k+1) j_k^2 = 1 (GSL sf_bessel_jl is called at
// integrals.c:309-316)
__device__ double sph_jl(int l, double x)
{
    const double ax = fabs(x);
    if (ax == 0.0) return l == 0 ? 1.0 : 0.0;
    if (ax < 0.5) {
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
    double s, c;
    sincos(x, &s, &c);
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
    const int L = l + 30 + (int)ax;
    double fp = 0.0, fc = 1.0, fl = 0.0, norm = 0.0, f0 = 0.0, f1 = 0.0;
    for (int k = L; k >= 0; --k) {
        norm += (2.0 * k + 1.0) * fc * fc;
        if (k == l) fl = fc;
        if (k == 1) f1 = fc;
        if (k == 0) {
            f0 = fc;
            break;
        }
        const double fm = (2.0 * k + 1.0) / x * fc - fp;
        fp = fc;
        fc = fm;
        if (fabs(fc) > 1e250) {
            fc *= 1e-250; fp *= 1e-250; fl *= 1e-250; f1 *= 1e-250;
            norm *= 1e-250; norm *= 1e-250;
        }
    }
    double scale = 1.0 / sqrt(norm);
    const double ref = fabs(j0) > fabs(j1) ? j0 : j1;
    const double fref = fabs(j0) > fabs(j1) ? f0 : f1;
    if ((ref < 0) != (fref * scale < 0)) scale = -scale;
    return fl * scale;
}

// cylindrical Bessel J_m(x), integer m >= 0, x >= 0 (the half-integer tables' small-r
// integrand calls GSL's sf_bessel_Jn, integrals.c:296-306): power series below 2,
// Hankel's asymptotic expansion from 25 on (the terms of P and Q fall below 1e-17
// there for the orders used), otherwise Miller's downward recurrence normalised with
// J_0 + 2 sum_k J_2k = 1, started 30 orders above max(m, x)
__device__ double cyl_jn(int m, double x)
{
    if (x == 0.0) return m == 0 ? 1.0 : 0.0;
    if (x < 2.0) {
        double pre = 1.0;
        for (int i = 1; i <= m; ++i) pre *= 0.5 * x / i;
        const double y = -0.25 * x * x;
        double term = 1.0, sum = 1.0;
        for (int k = 1; k < 40; ++k) {
            term *= y / (k * (double)(k + m));
            sum += term;
            if (fabs(term) < 1e-18 * fabs(sum)) break;
        }
        return pre * sum;
    }
    const double pi = 3.14159265358979323846;
    if (x >= 25.0) {
        const double mu = 4.0 * m * m;
        double P = 1.0, Q = 0.0, ak = 1.0;
        for (int k = 1; k < 60; ++k) {
            ak *= (mu - (2.0 * k - 1.0) * (2.0 * k - 1.0)) / (k * 8.0 * x);
            if (k & 1)
                Q += (((k - 1) / 2) & 1) ? -ak : ak;
            else
                P += ((k / 2) & 1) ? -ak : ak;
            if (fabs(ak) < 1e-17) break;
        }
        double sn, cs;
        sincos(x - (0.5 * m + 0.25) * pi, &sn, &cs);
        return sqrt(2.0 / (pi * x)) * (P * cs - Q * sn);
    }
    const int top = m > (int)x ? m : (int)x;
    const int N = 2 * ((top + 30) / 2 + 1);
    double fp = 0.0, fc = 1e-30, fm = 0.0, norm = 0.0;
    for (int k = N; k > 0; --k) {
        if (k == m) fm = fc;
        if ((k & 1) == 0) norm += 2.0 * fc;
        const double fn = 2.0 * k / x * fc - fp;
        fp = fc;
        fc = fn;
        if (fabs(fc) > 1e250) {
            fc *= 1e-250;
            fp *= 1e-250;
            fm *= 1e-250;
            norm *= 1e-250;
        }
    }
    if (m == 0) fm = fc;
    norm += fc;
    return fm / norm;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// integrands over k of the Akima spline P(k) of the input samples
enum {
    SR_BESSEL = 0,    // k^2 P j_l(k r) / (k r)^n              (integrals.c:280-325)
    SR_PREFACTOR = 1, // k^(2+l-n) P                            (integrals.c:238-269)
    SR_RENORM0 = 2,   // P (1 - j_0(k r)^2) / k^2                (integrals.c:389-398)
    SR_RENORM = 3     // P j_0(k r) j_0(k r2) / k^2              (integrals.c:405-417)
};

struct sr_unit {
    const double *k, *p;   // input samples of the spectrum
    int n_in;
    int n, l;
    double r, r2;
    int kind;
    // half-integer table (n, l both half-integers, integrals.c:186-207): nd, ld are
    // n and l, and the kernel is sqrt(pi/2) J_{l+1/2}(kr) / (kr)^(n+1/2) = j_l(kr)/(kr)^n
    int half;
    double nd, ld;
};

__device__ __forceinline__ double integrand(const sr_unit &u, double k)
{
    const double pk = ce::akima_general(u.k, u.p, u.n_in, k);
    const double sqrt_pi_2 = 1.25331413731550025121;
    if (u.half && u.kind == SR_PREFACTOR) return sqrt_pi_2 * pow(k, 2.0 + u.ld - u.nd) * pk;
    if (u.half && u.kind == SR_BESSEL) {
        const double x = k * u.r;
        return k * k * pk * sqrt_pi_2 * cyl_jn((int)(u.ld + 0.5), x) / pow(x, u.nd + 0.5);
    }
    switch (u.kind) {
    case SR_PREFACTOR: return pow(k, (double)(2 + u.l - u.n)) * pk;
    case SR_RENORM0: {
        const double j0 = sph_jl(0, k * u.r);
        return pk * (1. - j0 * j0) / k / k;
    }
    case SR_RENORM: {
        // the product of the two Bessel factors first: symmetric in (r, r2) bit for bit
        const double jj = sph_jl(0, k * u.r) * sph_jl(0, k * u.r2);
        return pk * jj / k / k;
    }
    default: {
        const double x = k * u.r;
        return k * k * pk * sph_jl(u.l, x) / pow(x, (double)u.n);
    }
    }
}

struct qk_out {
    double result, abserr, resabs, resasc;
};

// QUADPACK qk61 on [a, b]: lane j < 61 evaluates node j of the rule (j < 30: the left
// half, j = 30: the centre, j > 30: the right half)
__device__ qk_out qk61(const sr_unit &u, double a, double b, int lane)
{
    const double center = 0.5 * (a + b), half = 0.5 * (b - a);
    double f = 0.0, wk = 0.0, wg = 0.0;
    if (lane < 61) {
        const int j = lane <= 30 ? lane : 60 - lane;   // QUADPACK index of the node
        const double x = lane < 30 ? center - half * c_xgk[j] : (lane == 30 ? center : center + half * c_xgk[j]);
        f = integrand(u, x);
        wk = c_wgk[j];
        wg = (j & 1) ? c_wg[j >> 1] : 0.0;
    }
    const double rk = wave_sum(wk * f);
    const double rg = wave_sum(wg * f);
    const double rabs = wave_sum(wk * fabs(f));
    const double mean = 0.5 * rk;
    const double rasc = wave_sum(wk * fabs(f - mean));
    qk_out o;
    double err = fabs((rk - rg) * half);
    o.result = rk * half;
    o.resabs = rabs * fabs(half);
    o.resasc = rasc * fabs(half);
    if (o.resasc != 0.0 && err != 0.0) {
        const double scale = pow(200.0 * err / o.resasc, 1.5);
        err = scale < 1.0 ? o.resasc * scale : o.resasc;
    }
    if (o.resabs > DBL_MIN / (50.0 * DBL_EPSILON)) {
        const double min_err = 50.0 * DBL_EPSILON * o.resabs;
        if (min_err > err) err = min_err;
    }
    o.abserr = err;
    return o;
}

struct sr_args {
    int nspec, ntab;
    const ce_spec_dev *spec;
    const ce_grid_dev *grids;
    const ce_table_desc *tabs;
    const double *k, *p;
    double epsrel;
    double *out;      // [nspec][ntab][1 + CE_MIN_SEP_LEN]
    int *status;      // same layout: 0 converged, 1 interval limit, 2 roundoff, 3 bad integrand
};

// QUADPACK QAG (epsabs 0, the 61-point rule) of u over [lo, hi] by one wavefront, the
// interval list in LDS (cap CE_QAG_CAP); every lane returns the same sum.  Status: 0
// converged, 1 interval limit, 2 roundoff, 3 bad integrand behaviour.
__device__ int qag_wave(const sr_unit &u, double lo, double hi, double epsrel, int lane, double *ia, double *ib,
                        double *ir, double *ie, double &result, int *nint = nullptr)
{
    if (nint) *nint = 1;
    qk_out q0 = qk61(u, lo, hi, lane);
    int st = 0;
    result = q0.result;
    double tol = epsrel * fabs(q0.result);
    const double round_off = 50.0 * DBL_EPSILON * q0.resabs;
    const bool done0 = (q0.abserr <= round_off && q0.abserr > tol) ||
                       (q0.abserr <= tol && q0.abserr != q0.resasc) || q0.abserr == 0.0;
    if (q0.abserr <= round_off && q0.abserr > tol) st = 2;
    if (done0) return st;
    if (lane == 0) {
        ia[0] = lo; ib[0] = hi; ir[0] = q0.result; ie[0] = q0.abserr;
    }
    __syncthreads();
    int size = 1, imax = 0, iter = 1, rt1 = 0, rt2 = 0;
    double area = q0.result, errsum = q0.abserr;
    do {
        const double a1 = ia[imax], b2 = ib[imax];
        const double b1 = 0.5 * (a1 + b2), a2 = b1;
        const double ri = ir[imax], ei = ie[imax];
        const qk_out q1 = qk61(u, a1, b1, lane);
        const qk_out q2 = qk61(u, a2, b2, lane);
        const double area12 = q1.result + q2.result, err12 = q1.abserr + q2.abserr;
        errsum += err12 - ei;
        area += area12 - ri;
        if (q1.resasc != q1.abserr && q2.resasc != q2.abserr) {
            const double delta = ri - area12;
            if (fabs(delta) <= 1.0e-5 * fabs(area12) && err12 >= 0.99 * ei) rt1++;
            if (iter >= 10 && err12 > ei) rt2++;
        }
        tol = epsrel * fabs(area);
        if (errsum > tol) {
            if (rt1 >= 6 || rt2 >= 20) st = 2;
            const double tmp = (1 + 100 * DBL_EPSILON) * (fabs(a2) + 1000 * DBL_MIN);
            if (fabs(a1) <= tmp && fabs(b2) <= tmp) st = 3;
        }
        __syncthreads();   // every lane has read the interval being replaced
        if (lane == 0) {
            // the refined interval keeps the half with the larger error
            const bool right = q2.abserr > q1.abserr;
            ia[imax] = right ? a2 : a1; ib[imax] = right ? b2 : b1;
            ir[imax] = right ? q2.result : q1.result; ie[imax] = right ? q2.abserr : q1.abserr;
            ia[size] = right ? a1 : a2; ib[size] = right ? b1 : b2;
            ir[size] = right ? q1.result : q2.result; ie[size] = right ? q1.abserr : q2.abserr;
        }
        __syncthreads();
        size++;
        // the interval with the largest error, first one on ties
        double be = -1.0;
        int bi = 0x7fffffff;
        for (int kk = lane; kk < size; kk += 64) {
            if (ie[kk] > be) { be = ie[kk]; bi = kk; }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double oe = __shfl_xor(be, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            if (oe > be || (oe == be && oi < bi)) { be = oe; bi = oi; }
        }
        imax = bi < size ? bi : 0;
        iter++;
        if (iter >= CE_QAG_CAP && errsum > tol && !st) st = 1;
    } while (iter < CE_QAG_CAP && !st && errsum > tol);
    double part = 0.0;
    for (int kk = lane; kk < size; kk += 64) part += ir[kk];
    result = wave_sum(part);
    if (nint) *nint = size;
    return st;
}

// one wavefront (block of 64) per (spectrum, table, node)
__global__ void __launch_bounds__(64) k_small_r(sr_args a)
{
    __shared__ double ia[CE_QAG_CAP], ib[CE_QAG_CAP], ir[CE_QAG_CAP], ie[CE_QAG_CAP];
    constexpr int NODES = 1 + CE_MIN_SEP_LEN;
    const int unit = blockIdx.x;
    const int lane = threadIdx.x;
    const int node = unit % NODES;
    const int t = (unit / NODES) % a.ntab;
    const int s = unit / (NODES * a.ntab);
    const ce_spec_dev sp = a.spec[s];
    const ce_grid_dev &g = a.grids[sp.grid];
    sr_unit u;
    u.k = a.k + sp.in_off;
    u.p = a.p + sp.in_off;
    u.n_in = sp.n_in;
    u.n = (int)a.tabs[t].n;
    u.l = (int)a.tabs[t].l;
    u.nd = a.tabs[t].n;
    u.ld = a.tabs[t].l;
    u.half = u.ld != floor(u.ld);
    u.kind = node == 0 ? SR_PREFACTOR : SR_BESSEL;
    u.r = node == 0 ? 0.0 : c_min_sep[node - 1];
    u.r2 = 0.0;
    const int64_t o = (int64_t)unit;
    if (node == 0 && (u.half ? u.nd < u.ld : u.n < u.l)) {   // I_l^n(0) = 0 for n < l
        if (lane == 0) {
            a.out[o] = 0.0;
            a.status[o] = 0;
        }
        return;
    }
    double result;
    int st = qag_wave(u, g.kmin, g.kmax, a.epsrel, lane, ia, ib, ir, ie, result);
    if (lane == 0) {
        const double pi = 3.14159265358979323846;
        double v = result / 2. / pi / pi;
        if (node == 0 && u.half) {
            // 2^(l+1/2) Gamma(l+3/2) = 2^m m!, m = l + 1/2 (integrals.c:215-231)
            const int m = (int)(u.ld + 0.5);
            double c = 1.0;
            for (int i = 1; i <= m; ++i) c *= 2.0 * i;
            v /= c;
        } else if (node == 0) {
            double df = 1.0;   // (2l+1)!! = 2^(l+1/2) Gamma(l+3/2) / sqrt(pi/2)
            for (int i = 3; i <= 2 * u.l + 1; i += 2) df *= i;
            v /= df;
        } else if (u.half ? u.nd >= u.ld : u.n >= u.l) {
            v *= u.half ? pow(u.r, u.nd - u.ld) : pow(u.r, (double)(u.n - u.l));
        }
        if (!isfinite(v) && st == 0) st = 3;
        a.out[o] = v;
        a.status[o] = st;
    }
}

struct div_args {
    const double *k, *p;   // input samples of the spectrum
    int n_in;
    double kmin, kmax, epsrel;
    const double *sep;     // [nsep]
    int nsep;
    const double *chi;     // [nchi]
    int nchi;
    double *i40, *r0;      // [nsep] each
    double *ren;           // [nchi][nchi]: (chi[j], chi[i]) at i nchi + j
    int *status;           // [2 nsep + nchi^2]: the QAG status, 4 for a non-finite value
    int *nint;             // same layout: final interval count, or null
};

// the divergent table I_0^4 and its renormalisations (integrals.c:772-909), one
// wavefront per integral: units [0, nsep) I_0^4(sep) sep^4, [nsep, 2 nsep) the r = 0
// renormalisation at sep (times sep^4 as well, integrals.c:351-356, n = 4, l = 0), then
// the nchi^2 grid of int P j_0 j_0 / k^2; all divided by 2 pi^2
__global__ void __launch_bounds__(64) k_divergent(div_args a)
{
    __shared__ double ia[CE_QAG_CAP], ib[CE_QAG_CAP], ir[CE_QAG_CAP], ie[CE_QAG_CAP];
    const int unit = blockIdx.x;
    const int lane = threadIdx.x;
    sr_unit u;
    u.k = a.k;
    u.p = a.p;
    u.n_in = a.n_in;
    u.n = 4;
    u.l = 0;
    u.half = 0;
    u.nd = 4.0;
    u.ld = 0.0;
    u.r2 = 0.0;
    if (unit < 2 * a.nsep) {
        u.kind = unit < a.nsep ? SR_BESSEL : SR_RENORM0;
        u.r = a.sep[unit < a.nsep ? unit : unit - a.nsep];
    } else {
        const int q = unit - 2 * a.nsep;
        u.kind = SR_RENORM;
        u.r = a.chi[q % a.nchi];
        u.r2 = a.chi[q / a.nchi];
    }
    double result;
    int nint;
    int st = qag_wave(u, a.kmin, a.kmax, a.epsrel, lane, ia, ib, ir, ie, result, &nint);
    if (lane == 0) {
        const double pi = 3.14159265358979323846;
        if (a.nint) a.nint[unit] = nint;
        double v = result;
        if (u.kind != SR_RENORM && u.r > 0) v *= pow(u.r, 4.0);
        v = v / 2. / pi / pi;
        if (!isfinite(v)) st = 4;
        if (unit < a.nsep) a.i40[unit] = v;
        else if (unit < 2 * a.nsep) a.r0[unit - a.nsep] = v;
        else a.ren[unit - 2 * a.nsep] = v;
        a.status[unit] = st;
    }
}

}  // namespace

extern "C" int ce_launch_small_r(void *stream, int nspec, int ntab, const ce_spec_dev *d_spec,
                                 const ce_grid_dev *d_grids, const ce_table_desc *d_tabs, const double *d_k,
                                 const double *d_pk, double epsrel, double *d_out, int *d_status)
{
    static const double min_sep[CE_MIN_SEP_LEN] = {
        1E-6, 1E-5, 1E-4, 1E-3, 2E-3, 5E-3, 7E-3, 8E-3, 9E-3, 1E-2, 1.1E-2, 1.2E-2, 1.25E-2, 1.3E-2,
        1.35E-2, 1.5E-2, 2E-2, 2.5E-2, 3E-2, 3.5E-2, 5E-2, 7E-2, 8E-2, 9E-2, 1E-1, 1.1E-1, 1.2E-1,
        1.25E-1, 1.3E-1, 1.35E-1, 1.5E-1, 1.75E-1, 2E-1, 2.5E-1, 3E-1, 3.5E-1, 4E-1, 5E-1, 6E-1,
        7E-1, 8E-1, 9E-1};
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_min_sep), min_sep, sizeof(min_sep), 0,
                                          hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_xgk), GK61K_X, sizeof(GK61K_X), 0, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_wgk), GK61K_W, sizeof(GK61K_W), 0, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_wg), GK61G_W, sizeof(GK61G_W), 0, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return (int)e;
    sr_args a;
    a.nspec = nspec; a.ntab = ntab; a.spec = d_spec; a.grids = d_grids; a.tabs = d_tabs;
    a.k = d_k; a.p = d_pk; a.epsrel = epsrel; a.out = d_out; a.status = d_status;
    const int units = nspec * ntab * (1 + CE_MIN_SEP_LEN);
    hipLaunchKernelGGL(k_small_r, dim3(units), dim3(64), 0, s, a);
    return (int)hipGetLastError();
}

extern "C" int ce_launch_divergent(void *stream, const double *d_k, const double *d_pk, int n_in, double kmin,
                                   double kmax, const double *d_sep, int nsep, const double *d_chi, int nchi,
                                   double epsrel, double *d_i40, double *d_r0, double *d_ren, int *d_status,
                                   int *d_nint)
{
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_xgk), GK61K_X, sizeof(GK61K_X), 0, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_wgk), GK61K_W, sizeof(GK61K_W), 0, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyToSymbolAsync(HIP_SYMBOL(c_wg), GK61G_W, sizeof(GK61G_W), 0, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return (int)e;
    div_args a;
    a.k = d_k; a.p = d_pk; a.n_in = n_in; a.kmin = kmin; a.kmax = kmax; a.epsrel = epsrel;
    a.sep = d_sep; a.nsep = nsep; a.chi = d_chi; a.nchi = nchi;
    a.i40 = d_i40; a.r0 = d_r0; a.ren = d_ren; a.status = d_status; a.nint = d_nint;
    const int units = 2 * nsep + nchi * nchi;
    if (units <= 0) return 0;
    hipLaunchKernelGGL(k_divergent, dim3(units), dim3(64), 0, s, a);
    return (int)hipGetLastError();
}
