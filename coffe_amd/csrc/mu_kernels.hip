// mu-quadrature weights of the local multipoles (gfx950, FP64).
//
// The local correlation function is linear in the I_l^n tables evaluated at the
// separation (functions_nonintegrated, functions.c:89-1484):
//     xi(z, mu, r) = G(z, mu, r) * sum_t c_t(z, mu, r) I_t(r),
// with G = D1(z1) D1(z2) in full sky and D1(z)^2 in flat sky (functions.c:1457-1467),
// so the multipole (2l+1) int_0^1 dx xi(2x-1) P_l(2x-1)  (signal.c:83-111, 521-542)
// equals sum_t A[z][l][t][r] I_t(r) with spectrum-independent weights
//     A[z][l][t][r] = (2l+1)/2 sum_q w_q P_l(mu_q) G c_t(mu_q)
// for a Gauss-Legendre rule (mu_q, w_q) on [-1, 1].  One wavefront computes one
// (z, r) point: each lane evaluates its share of the nodes, then the per-(l, t)
// partial sums are reduced across the wavefront with lane shuffles.
//
// Background functions are Akima splines of their tables (the default
// interp_method, parser.c:807), evaluated on the fly from the six-point stencil
// around the interval with GSL's end-slope extension.
#include <hip/hip_runtime.h>

#include <string.h>

#include "ce_internal.h"
#include "coffe_engine.h"
#include "ce_device.h"
#include "fullsky_forms.h"

#include <type_traits>

namespace {

__device__ __forceinline__ double legendre_p(int l, double x)
{
    if (l == 0) return 1.0;
    double p0 = 1.0, p1 = x;
    for (int k = 2; k <= l; ++k) {
        const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

__device__ __forceinline__ double sq(double x) { return x * x; }

#define CE_MU_MAXL 4
// Akima coefficient tables: z(chi); f, conformal H, G1, G2, D1, a on the z grid; bias1,
// bias2 on bias_z; s1, s2, f_evo1, f_evo2 on the magnification grid (the last five are
// built only for the d2 / g terms)
#define CE_MU_NCOEF 13
enum { K_ZCHI = 0, K_F, K_H, K_G1, K_G2, K_D1, K_B1, K_B2, K_A, K_S1, K_S2, K_E1, K_E2 };

// table slots: the 8 standard tables (integrals.c:599-608), the divergent I_0^4
// (integrals.c:772-834), then the accumulator of the renormalisation term
enum { S00 = 0, S02, S04, S11, S13, S20, S22, S31, S40, NSTD, SREN = NSTD, NACC };

struct mu_args {
    int nbg;
    const double *bg_z, *bg_chi, *bg_D1, *bg_f, *bg_cH, *bg_G1, *bg_G2, *bg_a;
    int nb;
    const double *bias_z, *bias1, *bias2;
    int nmb;                   // magnification / evolution biases (d2 / g terms)
    const double *mb_z, *s1, *s2, *fevo1, *fevo2;
    double H0h, Om;            // COFFE_H0 h (1/Mpc), Omega0_cdm + Omega0_baryon
    unsigned terms;            // COFFE_TERM_* mask
    unsigned dg;               // sources of the d2 / g terms: bit 0 d2, 1 g1, 2 g2, 3 g3
    int flatsky;
    int paired;                // point i is (zmean[i], sep[i]) instead of the z x r grid
    int nz;
    const double *zmean;
    int nl;
    int ls[CE_MAX_TABLES];
    int ntab;
    int slot[CE_MAX_TABLES + 1];   // slot of output table t (the 8 standard ones and I_0^4)
    int nsep;
    const double *sep;
    int nmu;
    const double *mu, *wmu;
    double *A;                 // [nz][nl][ntab][nsep] (paired: [npoint][nl][ntab])
    // the renormalisation grid of I_0^4 (integrals.c:848-907): values and the bicubic's
    // node derivatives on ren_chi x ren_chi, [b nren + a] at (chi_a, chi_b)
    int nren;
    const double *ren_chi, *ren_v, *ren_x, *ren_y, *ren_xy;
    double *ren_out;           // [nz][nl][nsep] (paired: [npoint][nl]) or null
    const double *cb[CE_MU_NCOEF], *cc[CE_MU_NCOEF], *cd_[CE_MU_NCOEF];
    // the JVP instance: the tangents of chi, f, conformal H, G1, G2 and D1 on the z grid
    // along one direction, the tangents of their Akima coefficients (null: zero), and
    // the tangent weights, laid out as A
    const double *tg_chi, *tg_f, *tg_cH, *tg_G1, *tg_G2, *tg_D1;
    const double *cbt[CE_MU_NCOEF], *cct[CE_MU_NCOEF], *cdt[CE_MU_NCOEF];
    double *At;
};

// y(xq) in interval j from precomputed Akima coefficients: the value akima_at returns
__device__ __forceinline__ double akima_coef(const double *__restrict__ x, const double *__restrict__ y,
                                             const double *__restrict__ b, const double *__restrict__ c,
                                             const double *__restrict__ d, int j, double xq)
{
    const double delx = xq - x[j];
    return y[j] + delx * (b[j] + delx * (c[j] + d[j] * delx));
}

struct coef_args {
    const double *x[CE_MU_NCOEF], *y[CE_MU_NCOEF];
    const double *xt[CE_MU_NCOEF], *yt[CE_MU_NCOEF];   // node tangents (dual instance; null: 0)
    int n[CE_MU_NCOEF];
    double *b[CE_MU_NCOEF], *c[CE_MU_NCOEF], *d[CE_MU_NCOEF];
    double *bt[CE_MU_NCOEF], *ct[CE_MU_NCOEF], *dt[CE_MU_NCOEF];   // their tangents (dual instance)
};

__device__ __forceinline__ double tangent_of(double) { return 0.0; }
__device__ __forceinline__ double tangent_of(ce::dd x) { return x.t; }

// Akima (GSL, non-rounded ends) coefficients of every interval of the tables: one thread
// per (interval, table), ce::akima_coefs; T = dd also writes their tangents along the
// nodes' tangents
template <typename T>
__global__ void __launch_bounds__(256) k_akima_coef(coef_args a)
{
    const int t = blockIdx.y;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = a.n[t];
    if (j >= n - 1) return;
    const ce::akima_cf<T> cf = ce::akima_coefs<T>(a.x[t], a.xt[t], a.y[t], a.yt[t], n, j);
    a.b[t][j] = ce::val(cf.b);
    a.c[t][j] = ce::val(cf.c);
    a.d[t][j] = ce::val(cf.d);
    if (a.bt[t]) {
        a.bt[t][j] = tangent_of(cf.b);
        a.ct[t][j] = tangent_of(cf.c);
        a.dt[t][j] = tangent_of(cf.d);
    }
}

// a table value by its Akima coefficients in interval j (akima_coef); the dual instance
// also carries the tangents of the nodes (xt, yt) and of the coefficients of table K
template <typename T>
__device__ __forceinline__ T lookup(const mu_args &a, int K, const double *__restrict__ x,
                                    const double *__restrict__ xt, const double *__restrict__ y,
                                    const double *__restrict__ yt, int j, T q)
{
    if constexpr (std::is_same<T, double>::value) {
        return akima_coef(x, y, a.cb[K], a.cc[K], a.cd_[K], j, q);
    } else {
        auto dual = [&](const double *v, const double *vt) { return T(v[j], vt ? vt[j] : 0.0); };
        const T delx = q - dual(x, xt);
        return dual(y, yt) + delx * (dual(a.cb[K], a.cbt[K]) + delx * (dual(a.cc[K], a.cct[K]) +
                                                                       dual(a.cd_[K], a.cdt[K]) * delx));
    }
}

// the background at the end point chi of population p (1 or 2): z(chi) and f, conformal
// H, G_p, D1 at that z (one interval search on the chi table, started from a linear
// extrapolation of the mean's interval, one on the z grid shared by the lookups), the
// galaxy bias b_p, and for the d2 / g terms a(z), s_p and f_evo,p
template <typename T>
struct endpoint_t {
    T z, f, H, G, D1, b, a, s, fevo;
};
using endpoint = endpoint_t<double>;

template <typename T, bool DG>
__device__ __forceinline__ endpoint_t<T> background_at(const mu_args &a, T chi, int p, int jbar, double chib,
                                                       double dchi)
{
    endpoint_t<T> e;
    const int n = a.nbg;
    const double nan = __builtin_nan("");
    e.a = e.s = e.fevo = 0.0;
    const double chiv = ce::val(chi);
    if (!(chiv >= a.bg_chi[0] && chiv <= a.bg_chi[n - 1])) {
        e.z = e.f = e.H = e.G = e.D1 = e.b = e.a = e.s = e.fevo = nan;
        return e;
    }
    const int jc = ce::interval_from(a.bg_chi, n, chiv, ce::clamp_index(jbar + floor((chiv - chib) / dchi), n));
    // z(chi): the chi nodes move with the background, the z nodes do not
    e.z = lookup<T>(a, K_ZCHI, a.bg_chi, a.tg_chi, a.bg_z, nullptr, jc, chi);
    const double zv = ce::val(e.z);
    const int k = ce::interval_from(a.bg_z, n, zv, ce::interval_guess(a.bg_z, n, zv));
    e.f = lookup<T>(a, K_F, a.bg_z, nullptr, a.bg_f, a.tg_f, k, e.z);
    e.H = lookup<T>(a, K_H, a.bg_z, nullptr, a.bg_cH, a.tg_cH, k, e.z);
    const int kg = p == 1 ? K_G1 : K_G2;
    e.G = lookup<T>(a, kg, a.bg_z, nullptr, p == 1 ? a.bg_G1 : a.bg_G2, p == 1 ? a.tg_G1 : a.tg_G2, k, e.z);
    e.D1 = lookup<T>(a, K_D1, a.bg_z, nullptr, a.bg_D1, a.tg_D1, k, e.z);
    if (!(zv >= a.bias_z[0] && zv <= a.bias_z[a.nb - 1])) {
        e.b = nan;
    } else {
        const int kb = ce::interval_from(a.bias_z, a.nb, zv, ce::interval_guess(a.bias_z, a.nb, zv));
        const int kt = p == 1 ? K_B1 : K_B2;
        e.b = lookup<T>(a, kt, a.bias_z, nullptr, p == 1 ? a.bias1 : a.bias2, nullptr, kb, e.z);
    }
    if constexpr (DG) {
        e.a = akima_coef(a.bg_z, a.bg_a, a.cb[K_A], a.cc[K_A], a.cd_[K_A], k, e.z);
        if (!(e.z >= a.mb_z[0] && e.z <= a.mb_z[a.nmb - 1])) {
            e.s = e.fevo = nan;
        } else {
            const int km = ce::interval_from(a.mb_z, a.nmb, e.z, ce::interval_guess(a.mb_z, a.nmb, e.z));
            const int ks = p == 1 ? K_S1 : K_S2, ke = p == 1 ? K_E1 : K_E2;
            e.s = akima_coef(a.mb_z, p == 1 ? a.s1 : a.s2, a.cb[ks], a.cc[ks], a.cd_[ks], km, e.z);
            e.fevo = akima_coef(a.mb_z, p == 1 ? a.fevo1 : a.fevo2, a.cb[ke], a.cc[ke], a.cd_[ke], km, e.z);
        }
    }
    return e;
}

// the d2 / g source amplitude of one end point: the d2 and g1-g3 contributions of
// functions.c:574-669, 856-903, 1144-1428 are sums of products of one factor per end
// point, S(1) S(2) for the auto and mutual terms and b, f, H f times S for the crosses:
//   S = [d2] (3 - f_evo) H^2 f - sum_g [g] (H0 h)^2 3 Omega_m / (2 a) {1 + G, 5 s - 2, f - 1}
__device__ __forceinline__ double dg_source(const mu_args &a, const endpoint &e)
{
    double S = 0.0;
    if (a.dg & 1u) S += (3 - e.fevo) * e.H * e.H * e.f;
    const double g0 = a.H0h * a.H0h * 3 * a.Om / 2 / e.a;
    if (a.dg & 2u) S -= g0 * (1 + e.G);
    if (a.dg & 4u) S -= g0 * (5 * e.s - 2);
    if (a.dg & 8u) S -= g0 * (e.f - 1);
    return S;
}

// the auto part of S(1) S(2): the sum over the chosen sources of each one's factor at
// the two end points (the d2-d2, g1-g1, g2-g2, g3-g3 terms of functions.c:574-669)
__device__ __forceinline__ double dg_auto(const mu_args &a, const endpoint &e1, const endpoint &e2)
{
    double D = 0.0;
    if (a.dg & 1u) D += (3 - e1.fevo) * e1.H * e1.H * e1.f * ((3 - e2.fevo) * e2.H * e2.H * e2.f);
    const double g1 = a.H0h * a.H0h * 3 * a.Om / 2 / e1.a, g2 = a.H0h * a.H0h * 3 * a.Om / 2 / e2.a;
    if (a.dg & 2u) D += g1 * (1 + e1.G) * (g2 * (1 + e2.G));
    if (a.dg & 4u) D += g1 * (5 * e1.s - 2) * (g2 * (5 * e2.s - 2));
    if (a.dg & 8u) D += g1 * (e1.f - 1) * (g2 * (e2.f - 1));
    return D;
}

// the renormalisation grid's bicubic (GSL interp2d bicubic: node derivatives of natural
// cubic splines, Hermite patches; coffe_interp_spline2d, common.c:712-719) at (x, y)
__device__ __forceinline__ double ren_bicubic(const mu_args &a, double x, double y)
{
    const int n = a.nren;
    if (!(x >= a.ren_chi[0] && x <= a.ren_chi[n - 1] && y >= a.ren_chi[0] && y <= a.ren_chi[n - 1]))
        return __builtin_nan("");
    const int i = ce::interval_from(a.ren_chi, n, x, ce::interval_guess(a.ren_chi, n, x));
    const int j = ce::interval_from(a.ren_chi, n, y, ce::interval_guess(a.ren_chi, n, y));
    const double dx = a.ren_chi[i + 1] - a.ren_chi[i], dy = a.ren_chi[j + 1] - a.ren_chi[j];
    const double t = (x - a.ren_chi[i]) / dx, u = (y - a.ren_chi[j]) / dy;
    const double ht[4] = {(1 + 2 * t) * (1 - t) * (1 - t), t * t * (3 - 2 * t), t * (1 - t) * (1 - t), t * t * (t - 1)};
    const double hu[4] = {(1 + 2 * u) * (1 - u) * (1 - u), u * u * (3 - 2 * u), u * (1 - u) * (1 - u), u * u * (u - 1)};
    double r = 0.0;
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int o = (j + q) * n + i + p;
            r += a.ren_v[o] * ht[p] * hu[q] + a.ren_x[o] * dx * ht[2 + p] * hu[q] + a.ren_y[o] * dy * ht[p] * hu[2 + q] +
                 a.ren_xy[o] * dx * dy * ht[2 + p] * hu[2 + q];
        }
    return r;
}

// coefficients c_s of the local terms at one (z, mu, r), times G = D1(z1) D1(z2) (full
// sky) or D1(z_mean)^2 (flat sky), functions.c:1457-1467; c[SREN] is minus the
// coefficient of the renormalisation grid's value at (chi1, chi2)
template <typename T, bool DG>
__device__ __forceinline__ void local_coefficients(const mu_args &a, double mu, double sep, T chi_mean, T fmean,
                                                   T Hmean, T bm1, T bm2, T Gm1, T Gm2, T Dmean, int jbar,
                                                   double chib, double dchi, T *c)
{
    for (int s = 0; s < (DG ? NACC : S40); ++s) c[s] = 0.0;
    const T chi1 = chi_mean - sep * mu / 2.;
    const T chi2 = chi_mean + sep * mu / 2.;
    // the flat-sky forms of the den / rsd / d1 terms use the mean's background only; the
    // full-sky ones and every d2 / g term the end points' (the reference reads them in
    // flat sky too), with the geometric coefficients of fullsky_forms.h (the reference's
    // polynomials in cos theta with their (chi/s)^n cancellations done exactly)
    const bool ends = !a.flatsky || DG;
    endpoint_t<T> e1, e2;
    ce_fullsky_geom_t<T> g;
    if (ends) {
        e1 = background_at<T, DG>(a, chi1, 1, jbar, chib, dchi);
        e2 = background_at<T, DG>(a, chi2, 2, jbar, chib, dchi);
        g = ce_fullsky_geometry(chi_mean, sep, mu);
    } else {
        e1.z = e1.f = e1.H = e1.G = e1.D1 = e1.b = e1.a = e1.s = e1.fevo = 0.0;
        e2 = e1;
    }
    const T f1 = e1.f, f2 = e2.f, H1 = e1.H, H2 = e2.H, b1 = e1.b, b2 = e2.b;
    const T G1 = e1.G, G2 = e2.G;
    const double P2 = legendre_p(2, mu), P4 = legendre_p(4, mu);
    if (a.terms & COFFE_TERM_DEN_DEN) {   // functions.c:129-229
        if (a.flatsky) c[S00] += bm1 * bm2;
        else c[S00] += b1 * b2;
    }
    if (a.terms & COFFE_TERM_RSD_RSD) {   // functions.c:230-362
        if (a.flatsky) {
            c[S00] += fmean * fmean / 5.;
            c[S02] += -4 * fmean * fmean / 7. * P2;
            c[S04] += 8 * fmean * fmean / 35. * P4;
        } else {
            c[S00] += f1 * f2 * g.rr0;
            c[S02] += f1 * f2 * g.rr2;
            c[S04] += f1 * f2 * g.rr4;
        }
    }
    if (a.terms & COFFE_TERM_D1_D1) {   // functions.c:363-573
        if (a.flatsky) {
            c[S20] += Hmean * Hmean * fmean * fmean * Gm1 * Gm2 / 3.;
            c[S22] += 2 * Hmean * Hmean * fmean * fmean * Gm1 * Gm2 * sq(sep) * P2 / 3.;
        } else {
            c[S20] += H1 * H2 * f1 * f2 * G1 * G2 * g.dd0;
            c[S22] += H1 * H2 * f1 * f2 * G1 * G2 * g.dd2;
        }
    }
    if (a.terms & COFFE_TERM_DEN_RSD) {   // functions.c:670-771
        if (a.flatsky) {
            c[S00] += bm1 * fmean / 3. + bm2 * fmean / 3.;
            c[S02] += -(2 * bm1 * fmean / 3. + 2 * bm2 * fmean / 3.) * P2;
        } else {
            c[S00] += b1 * f2 / 3. + b2 * f1 / 3.;
            c[S02] += -(b1 * f2 * g.dr_a + b2 * f1 * g.dr_b);
        }
    }
    if (a.terms & COFFE_TERM_DEN_D1) {   // functions.c:772-855
        if (a.flatsky) {
            c[S11] += (bm1 * fmean * Hmean * Gm2 * mu + bm2 * fmean * Hmean * Gm1 * (-mu)) * sep;
        } else {
            c[S11] += -(b1 * f2 * H2 * G2 * g.dd1_a + b2 * f1 * H1 * G1 * g.dd1_b);
        }
    }
    if (a.terms & COFFE_TERM_RSD_D1) {   // functions.c:904-1143
        if (a.flatsky) {
            const double P3p = legendre_p(3, mu), P3m = legendre_p(3, -mu);
            c[S11] += -3. / 5. * (fmean * fmean * Hmean * Gm2 * (-mu) + fmean * fmean * Hmean * Gm1 * mu) * sep;
            c[S13] += (fmean * fmean * Hmean * Gm2 * 2. / 5. * P3m + fmean * fmean * Hmean * Gm1 * 2. / 5. * P3p) * sep;
        } else {
            c[S11] += f1 * f2 * H2 * G2 * g.rd1_a + f2 * f1 * H1 * G1 * g.rd1_b;
            c[S13] += f1 * f2 * H2 * G2 * g.rd3_a + f2 * f1 * H1 * G1 * g.rd3_b;
        }
    }
    if constexpr (DG) {
        const double Sg1 = dg_source(a, e1), Sg2 = dg_source(a, e2);
        if (a.terms & COFFE_TERM_DG_DG) {     // I_0^4 minus its renormalisation at (chi1, chi2)
            double P = Sg1 * Sg2;
            if (a.terms & COFFE_TERM_DG_MUTUAL) P -= dg_auto(a, e1, e2);
            c[S40] += P;
            c[SREN] -= P * ren_bicubic(a, chi1, chi2);
        }
        if (a.terms & COFFE_TERM_DEN_DG) c[S20] += b1 * Sg2 + b2 * Sg1;
        if (a.terms & COFFE_TERM_RSD_DG) {
            c[S20] += (f1 * Sg2 + f2 * Sg1) / 3.;
            c[S22] += -(f1 * Sg2 * g.dr_b + f2 * Sg1 * g.dr_a) * sq(sep);
        }
        if (a.terms & COFFE_TERM_D1_DG) c[S31] += -(H1 * f1 * Sg2 * g.dd1_b + H2 * f2 * Sg1 * g.dd1_a);
    }
    const T G = a.flatsky ? Dmean * Dmean : e1.D1 * e2.D1;
    for (int s = 0; s < (DG ? NACC : S40); ++s) c[s] *= G;
}

// one wavefront per (z, r) point (paired: per (zmean[i], sep[i])); blockDim = 256 (four
// points).  The lanes take 64 mu nodes at a time; after each batch every (multipole,
// slot) sum is reduced across the wavefront and added into the wavefront's LDS
// accumulators, so that no per-lane partial sums stay live over the node loop.
// NL = the number of multipoles kept (>= a.nl) and DG = the d2 / g terms, template
// arguments so that absent multipoles and terms cost no registers
template <int NL, bool DG>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_mu_weights(mu_args a)
{
    constexpr int NC = DG ? NACC : S40;   // slots in use
    __shared__ double sacc[4][CE_MU_MAXL * NACC];
    const int lane = threadIdx.x & 63;
    const int point = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int npoint = a.paired ? a.nz : a.nz * a.nsep;
    if (point >= npoint) return;   // whole wavefronts exit together
    double *acc = sacc[threadIdx.x >> 6];
    if (lane < NL * NACC) acc[lane] = 0.0;
    const int iz = a.paired ? point : point / a.nsep, ir = a.paired ? point : point % a.nsep;
    const double zm = a.zmean[iz], sep = a.sep[ir];
    const double chi_mean = ce::akima_general(a.bg_z, a.bg_chi, a.nbg, zm);
    const double fmean = ce::akima_general(a.bg_z, a.bg_f, a.nbg, zm);
    const double Hmean = ce::akima_general(a.bg_z, a.bg_cH, a.nbg, zm);
    const double bm1 = ce::akima_general(a.bias_z, a.bias1, a.nb, zm), bm2 = ce::akima_general(a.bias_z, a.bias2, a.nb, zm);
    const double Gm1 = ce::akima_general(a.bg_z, a.bg_G1, a.nbg, zm), Gm2 = ce::akima_general(a.bg_z, a.bg_G2, a.nbg, zm);
    const double Dmean = ce::akima_general(a.bg_z, a.bg_D1, a.nbg, zm);
    // the mean's interval on the z grid (the chi table shares its indices) and the local
    // slope of chi, from which the end points' chi intervals are extrapolated
    const int jbar = ce::interval_bisect(a.bg_z, 0, a.nbg - 1, zm);
    const double chib = a.bg_chi[jbar], dchi = a.bg_chi[jbar + 1] - a.bg_chi[jbar];
    for (int q0 = 0; q0 < a.nmu; q0 += 64) {
        const int q = q0 + lane;
        double c[NACC];
        double mu = 0.0, w = 0.0;
        if (q < a.nmu) {
            mu = a.mu[q];
            w = a.wmu[q];
            local_coefficients<double, DG>(a, mu, sep, chi_mean, fmean, Hmean, bm1, bm2, Gm1, Gm2, Dmean, jbar,
                                           chib, dchi, c);
        } else {
#pragma unroll
            for (int s = 0; s < NC; ++s) c[s] = 0.0;
        }
#pragma unroll
        for (int il = 0; il < NL; ++il) {
            if (il >= a.nl) break;
            const double pl = w * legendre_p(a.ls[il], mu);
#pragma unroll
            for (int s = 0; s < NC; ++s) {
                double v = pl * c[s];
                for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
                if (lane == 0) acc[il * NACC + s] += v;
            }
        }
    }
    // lane 0 wrote the sums; the other lanes read them after the wavefront's LDS writes
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0): the LDS stores have completed
    if (lane < NL * NACC) {
        const int il = lane / NACC, s = lane % NACC;
        if (il < a.nl && s < NC) {
            const double v = (2 * a.ls[il] + 1) * 0.5 * acc[lane];
            if (s == SREN) {
                if (a.ren_out)
                    a.ren_out[a.paired ? (int64_t)point * a.nl + il : ((int64_t)iz * a.nl + il) * a.nsep + ir] = v;
            } else {
                for (int t = 0; t < a.ntab; ++t)
                    if (a.slot[t] == s)
                        a.A[a.paired ? ((int64_t)point * a.nl + il) * a.ntab + t
                                     : (((int64_t)iz * a.nl + il) * a.ntab + t) * a.nsep + ir] = v;
            }
        }
    }
}

// the forward-mode instance (density, RSD, Doppler-1; no d2 / g): the weights and their
// tangents along the background tables' tangents (chi and the z(chi) nodes, f, conformal
// H, G1, G2, D1), the same node loop with dual numbers and two accumulators per slot
template <int NL>
__global__ void __launch_bounds__(256) k_mu_weights_jvp(mu_args a)
{
    using ce::dd;
    __shared__ double sacc[4][2 * CE_MU_MAXL * S40];
    const int lane = threadIdx.x & 63;
    const int point = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int npoint = a.paired ? a.nz : a.nz * a.nsep;
    if (point >= npoint) return;   // whole wavefronts exit together
    double *acc = sacc[threadIdx.x >> 6];
    for (int i = lane; i < 2 * NL * S40; i += 64) acc[i] = 0.0;
    const int iz = a.paired ? point : point / a.nsep, ir = a.paired ? point : point % a.nsep;
    const double zm = a.zmean[iz], sep = a.sep[ir];
    const int n = a.nbg;
    const double nan = __builtin_nan("");
    const bool inside = zm >= a.bg_z[0] && zm <= a.bg_z[n - 1];
    const int jbar = inside ? ce::interval_bisect(a.bg_z, 0, n - 1, zm) : 0;
    // the mean's background as duals: Akima of the tables with their tangents at zm
    auto mean = [&](const double *y, const double *yt) -> dd {
        if (!inside) return dd(nan, nan);
        const ce::akima_cf<dd> cf = ce::akima_coefs<dd>(a.bg_z, nullptr, y, yt, n, jbar);
        const double delx = zm - a.bg_z[jbar];
        return cf.y0 + delx * (cf.b + delx * (cf.c + cf.d * delx));
    };
    const dd chi_mean = mean(a.bg_chi, a.tg_chi), fmean = mean(a.bg_f, a.tg_f), Hmean = mean(a.bg_cH, a.tg_cH);
    const dd Gm1 = mean(a.bg_G1, a.tg_G1), Gm2 = mean(a.bg_G2, a.tg_G2), Dmean = mean(a.bg_D1, a.tg_D1);
    const dd bm1 = ce::akima_general(a.bias_z, a.bias1, a.nb, zm), bm2 = ce::akima_general(a.bias_z, a.bias2, a.nb, zm);
    const double chib = a.bg_chi[jbar], dchi = a.bg_chi[jbar + 1] - a.bg_chi[jbar];
    for (int q0 = 0; q0 < a.nmu; q0 += 64) {
        const int q = q0 + lane;
        dd c[S40];
        double mu = 0.0, w = 0.0;
        if (q < a.nmu) {
            mu = a.mu[q];
            w = a.wmu[q];
            local_coefficients<dd, false>(a, mu, sep, chi_mean, fmean, Hmean, bm1, bm2, Gm1, Gm2, Dmean, jbar, chib,
                                          dchi, c);
        } else {
#pragma unroll
            for (int s = 0; s < S40; ++s) c[s] = 0.0;
        }
#pragma unroll
        for (int il = 0; il < NL; ++il) {
            if (il >= a.nl) break;
            const double pl = w * legendre_p(a.ls[il], mu);
#pragma unroll
            for (int s = 0; s < S40; ++s) {
                double v = pl * c[s].v, t = pl * c[s].t;
                for (int off = 32; off > 0; off >>= 1) {
                    v += __shfl_xor(v, off, 64);
                    t += __shfl_xor(t, off, 64);
                }
                if (lane == 0) {
                    acc[2 * (il * S40 + s)] += v;
                    acc[2 * (il * S40 + s) + 1] += t;
                }
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0): the LDS stores have completed
    if (lane < NL * S40) {
        const int il = lane / S40, s = lane % S40;
        if (il < a.nl) {
            const double f = (2 * a.ls[il] + 1) * 0.5;
            for (int t = 0; t < a.ntab; ++t)
                if (a.slot[t] == s) {
                    const int64_t o = a.paired ? ((int64_t)point * a.nl + il) * a.ntab + t
                                               : (((int64_t)iz * a.nl + il) * a.ntab + t) * a.nsep + ir;
                    a.A[o] = f * acc[2 * lane];
                    a.At[o] = f * acc[2 * lane + 1];
                }
        }
    }
}

}  // namespace

namespace {

// the Akima coefficient tables (coef: 3 per node of each table; with tangents a further 3
// per node of the six background tables that carry them) and the kernel arguments
void prepare_args(void *stream, const ce_mu_dev *in, const ce_mu_tan *tan, unsigned terms, unsigned dg, int flatsky,
                  int paired, int nz, const double *zmean, int nl, const int *ls, int ntab, const int *slot, int nsep,
                  const double *sep, int nmu, const double *mu, const double *wmu, double *A, double *ren_out,
                  double *coef, mu_args &a)
{
    memset(&a, 0, sizeof(a));
    coef_args ca;
    memset(&ca, 0, sizeof(ca));
    {
        const double *xs[CE_MU_NCOEF] = {in->bg_chi, in->bg_z, in->bg_z, in->bg_z, in->bg_z, in->bg_z, in->bias_z,
                                         in->bias_z, in->bg_z, in->mb_z, in->mb_z, in->mb_z, in->mb_z};
        const double *ys[CE_MU_NCOEF] = {in->bg_z, in->bg_f, in->bg_cH, in->bg_G1, in->bg_G2, in->bg_D1, in->bias1,
                                         in->bias2, in->bg_a, in->s1, in->s2, in->fevo1, in->fevo2};
        const int ns[CE_MU_NCOEF] = {in->nbg, in->nbg, in->nbg, in->nbg, in->nbg, in->nbg, in->nb, in->nb, in->nbg,
                                     in->nmb, in->nmb, in->nmb, in->nmb};
        const int ntabs = dg ? CE_MU_NCOEF : K_A;   // the d2 / g tables only when asked for
        double *p = coef;
        int nmax = 0;
        for (int t = 0; t < ntabs; ++t) {
            const int n = ns[t];
            ca.x[t] = xs[t]; ca.y[t] = ys[t]; ca.n[t] = n;
            ca.b[t] = p; ca.c[t] = p + n; ca.d[t] = p + 2 * n;
            a.cb[t] = ca.b[t]; a.cc[t] = ca.c[t]; a.cd_[t] = ca.d[t];
            p += 3 * n;
            if (n > nmax) nmax = n;
        }
        if (tan) {
            // the z(chi) table's nodes move with chi; the others' values with their tangents
            ca.xt[K_ZCHI] = tan->chi;
            ca.yt[K_F] = tan->f;
            ca.yt[K_H] = tan->cH;
            ca.yt[K_G1] = tan->G1;
            ca.yt[K_G2] = tan->G2;
            ca.yt[K_D1] = tan->D1;
            for (int t = K_ZCHI; t <= K_D1; ++t) {
                const int n = ns[t];
                ca.bt[t] = p; ca.ct[t] = p + n; ca.dt[t] = p + 2 * n;
                a.cbt[t] = ca.bt[t]; a.cct[t] = ca.ct[t]; a.cdt[t] = ca.dt[t];
                p += 3 * n;
            }
            hipLaunchKernelGGL(k_akima_coef<ce::dd>, dim3((nmax + 255) / 256, ntabs), dim3(256), 0,
                               (hipStream_t)stream, ca);
            a.tg_chi = tan->chi; a.tg_f = tan->f; a.tg_cH = tan->cH; a.tg_G1 = tan->G1; a.tg_G2 = tan->G2;
            a.tg_D1 = tan->D1;
        } else {
            hipLaunchKernelGGL(k_akima_coef<double>, dim3((nmax + 255) / 256, ntabs), dim3(256), 0,
                               (hipStream_t)stream, ca);
        }
    }
    a.nbg = in->nbg; a.bg_z = in->bg_z; a.bg_chi = in->bg_chi; a.bg_D1 = in->bg_D1; a.bg_f = in->bg_f;
    a.bg_cH = in->bg_cH; a.bg_G1 = in->bg_G1; a.bg_G2 = in->bg_G2; a.bg_a = in->bg_a;
    a.nb = in->nb; a.bias_z = in->bias_z; a.bias1 = in->bias1; a.bias2 = in->bias2;
    a.nmb = in->nmb; a.mb_z = in->mb_z; a.s1 = in->s1; a.s2 = in->s2; a.fevo1 = in->fevo1; a.fevo2 = in->fevo2;
    a.H0h = in->H0h; a.Om = in->Om;
    a.nren = in->nren; a.ren_chi = in->ren_chi; a.ren_v = in->ren_v; a.ren_x = in->ren_x; a.ren_y = in->ren_y;
    a.ren_xy = in->ren_xy;
    a.terms = terms; a.dg = dg; a.flatsky = flatsky; a.paired = paired;
    a.nz = nz; a.zmean = zmean; a.nl = nl;
    for (int i = 0; i < nl; ++i) a.ls[i] = ls[i];
    a.ntab = ntab;
    for (int t = 0; t < ntab; ++t) a.slot[t] = slot[t];
    a.nsep = nsep; a.sep = sep; a.nmu = nmu; a.mu = mu; a.wmu = wmu; a.A = A; a.ren_out = ren_out;
}

}  // namespace

extern "C" int ce_launch_mu_weights_jvp(void *stream, const ce_mu_dev *in, const ce_mu_tan *tan, unsigned terms,
                                        int flatsky, int paired, int nz, const double *zmean, int nl, const int *ls,
                                        int ntab, const int *slot, int nsep, const double *sep, int nmu,
                                        const double *mu, const double *wmu, double *A, double *At, double *coef)
{
    if (nl > CE_MU_MAXL || ntab > CE_MAX_TABLES || !tan) return (int)hipErrorInvalidValue;
    mu_args a;
    prepare_args(stream, in, tan, terms, 0, flatsky, paired, nz, zmean, nl, ls, ntab, slot, nsep, sep, nmu, mu, wmu,
                 A, nullptr, coef, a);
    a.At = At;
    const int npoint = paired ? nz : nz * nsep;
    const int blocks = (npoint + 3) / 4;
    hipStream_t s = (hipStream_t)stream;
    if (nl <= 1) hipLaunchKernelGGL((k_mu_weights_jvp<1>), dim3(blocks), dim3(256), 0, s, a);
    else if (nl == 2) hipLaunchKernelGGL((k_mu_weights_jvp<2>), dim3(blocks), dim3(256), 0, s, a);
    else if (nl == 3) hipLaunchKernelGGL((k_mu_weights_jvp<3>), dim3(blocks), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_mu_weights_jvp<CE_MU_MAXL>), dim3(blocks), dim3(256), 0, s, a);
    return (int)hipGetLastError();
}

extern "C" int ce_launch_mu_weights(void *stream, const ce_mu_dev *in, unsigned terms, unsigned dg, int flatsky,
                                    int paired, int nz, const double *zmean, int nl, const int *ls, int ntab,
                                    const int *slot, int nsep, const double *sep, int nmu, const double *mu,
                                    const double *wmu, double *A, double *ren_out, double *coef)
{
    if (nl > CE_MU_MAXL || ntab > CE_MAX_TABLES + 1) return (int)hipErrorInvalidValue;
    mu_args a;
    prepare_args(stream, in, nullptr, terms, dg, flatsky, paired, nz, zmean, nl, ls, ntab, slot, nsep, sep, nmu, mu,
                 wmu, A, ren_out, coef, a);
    const int npoint = paired ? nz : nz * nsep;
    const int blocks = (npoint + 3) / 4;
    hipStream_t s = (hipStream_t)stream;
    if (dg) {
        if (nl <= 1) hipLaunchKernelGGL((k_mu_weights<1, true>), dim3(blocks), dim3(256), 0, s, a);
        else if (nl == 2) hipLaunchKernelGGL((k_mu_weights<2, true>), dim3(blocks), dim3(256), 0, s, a);
        else if (nl == 3) hipLaunchKernelGGL((k_mu_weights<3, true>), dim3(blocks), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_mu_weights<CE_MU_MAXL, true>), dim3(blocks), dim3(256), 0, s, a);
    } else {
        if (nl <= 1) hipLaunchKernelGGL((k_mu_weights<1, false>), dim3(blocks), dim3(256), 0, s, a);
        else if (nl == 2) hipLaunchKernelGGL((k_mu_weights<2, false>), dim3(blocks), dim3(256), 0, s, a);
        else if (nl == 3) hipLaunchKernelGGL((k_mu_weights<3, false>), dim3(blocks), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((k_mu_weights<CE_MU_MAXL, false>), dim3(blocks), dim3(256), 0, s, a);
    }
    return (int)hipGetLastError();
}
