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

#include "ce_internal.h"
#include "coffe_engine.h"
#include "ce_device.h"
#include "fullsky_forms.h"

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
#define CE_MU_NCOEF 8

// standard table slots (integrals.c:599-608)
enum { S00 = 0, S02, S04, S11, S13, S20, S22, S31, NSTD };

struct mu_args {
    int nbg;
    const double *bg_z, *bg_chi, *bg_D1, *bg_f, *bg_cH, *bg_G1, *bg_G2;
    int nb;
    const double *bias_z, *bias1, *bias2;
    unsigned terms;            // COFFE_TERM_* mask
    int flatsky;
    int nz;
    const double *zmean;
    int nl;
    int ls[CE_MAX_TABLES];
    int ntab;
    int slot[CE_MAX_TABLES];   // standard slot of output table t
    int nsep;
    const double *sep;
    int nmu;
    const double *mu, *wmu;
    double *A;                 // [nz][nl][ntab][nsep]
    // Akima coefficients (b, c, d per interval) of the end-point tables: z(chi), then
    // f, conformal H, G1, G2, D1 on the z grid, then bias1, bias2 on bias_z
    const double *cb[CE_MU_NCOEF], *cc[CE_MU_NCOEF], *cd_[CE_MU_NCOEF];
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
    int n[CE_MU_NCOEF];
    double *b[CE_MU_NCOEF], *c[CE_MU_NCOEF], *d[CE_MU_NCOEF];
};

// Akima (GSL, non-rounded ends) coefficients of every interval of the tables: one thread
// per (interval, table); the same slopes and formulas as akima_at
__global__ void __launch_bounds__(256) k_akima_coef(coef_args a)
{
    const int t = blockIdx.y;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = a.n[t];
    if (j >= n - 1) return;
    const double *x = a.x[t], *y = a.y[t];
    auto slope = [&](int i) -> double {
        if (i >= 0 && i <= n - 2) return (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
        const double m0 = (y[1] - y[0]) / (x[1] - x[0]);
        const double m1 = (y[2] - y[1]) / (x[2] - x[1]);
        const double mn2 = (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]);
        const double mn3 = (y[n - 2] - y[n - 3]) / (x[n - 2] - x[n - 3]);
        if (i == -1) return 2.0 * m0 - m1;
        if (i == -2) return 3.0 * m0 - 2.0 * m1;
        if (i == n - 1) return 2.0 * mn2 - mn3;
        return 3.0 * mn2 - 2.0 * mn3;
    };
    const double mm2 = slope(j - 2), mm1 = slope(j - 1), m0 = slope(j), mp1 = slope(j + 1), mp2 = slope(j + 2);
    const double ne = fabs(mp1 - m0) + fabs(mm1 - mm2);
    double b, c, d;
    if (ne == 0.0) {
        b = m0; c = 0.0; d = 0.0;
    } else {
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
        b = (1.0 - alpha) * mm1 + alpha * m0;
        c = (3.0 * m0 - 2.0 * b - tn) / h;
        d = (b + tn - 2.0 * m0) / (h * h);
    }
    a.b[t][j] = b;
    a.c[t][j] = c;
    a.d[t][j] = d;
}

// coefficients c_s (standard slots) of the local terms at one (z, mu, r), times G
// the background at the end point chi of a pair: z(chi) and f, conformal H, G, D1 at
// that z (one interval search on the chi table, started from a linear extrapolation of
// the mean's interval, and one on the z grid shared by the four lookups), galaxy bias
struct endpoint {
    double z, f, H, G1, G2, D1, b1, b2;
};

__device__ __forceinline__ endpoint background_at(const mu_args &a, double chi, int jbar, double chib, double dchi)
{
    endpoint e;
    const int n = a.nbg;
    if (!(chi >= a.bg_chi[0] && chi <= a.bg_chi[n - 1])) {
        e.z = e.f = e.H = e.G1 = e.G2 = e.D1 = e.b1 = e.b2 = __builtin_nan("");
        return e;
    }
    const int jc = ce::interval_from(a.bg_chi, n, chi, ce::clamp_index(jbar + floor((chi - chib) / dchi), n));
    e.z = akima_coef(a.bg_chi, a.bg_z, a.cb[0], a.cc[0], a.cd_[0], jc, chi);
    const int k = ce::interval_from(a.bg_z, n, e.z, ce::interval_guess(a.bg_z, n, e.z));
    e.f = akima_coef(a.bg_z, a.bg_f, a.cb[1], a.cc[1], a.cd_[1], k, e.z);
    e.H = akima_coef(a.bg_z, a.bg_cH, a.cb[2], a.cc[2], a.cd_[2], k, e.z);
    e.G1 = akima_coef(a.bg_z, a.bg_G1, a.cb[3], a.cc[3], a.cd_[3], k, e.z);
    e.G2 = akima_coef(a.bg_z, a.bg_G2, a.cb[4], a.cc[4], a.cd_[4], k, e.z);
    e.D1 = akima_coef(a.bg_z, a.bg_D1, a.cb[5], a.cc[5], a.cd_[5], k, e.z);
    if (!(e.z >= a.bias_z[0] && e.z <= a.bias_z[a.nb - 1])) {
        e.b1 = e.b2 = __builtin_nan("");
    } else {
        const int kb = ce::interval_from(a.bias_z, a.nb, e.z, ce::interval_guess(a.bias_z, a.nb, e.z));
        e.b1 = akima_coef(a.bias_z, a.bias1, a.cb[6], a.cc[6], a.cd_[6], kb, e.z);
        e.b2 = akima_coef(a.bias_z, a.bias2, a.cb[7], a.cc[7], a.cd_[7], kb, e.z);
    }
    return e;
}

__device__ __forceinline__ void local_coefficients(const mu_args &a, double zm, double mu, double sep, double chi_mean,
                                   double fmean, double Hmean, double bm1, double bm2, double Gm1, double Gm2,
                                   double Dmean, int jbar, double chib, double dchi, double *c)
{
    for (int s = 0; s < NSTD; ++s) c[s] = 0.0;
    const double chi1 = chi_mean - sep * mu / 2.;
    const double chi2 = chi_mean + sep * mu / 2.;
    // the flat-sky forms use the mean's background only; the full-sky ones the end
    // points' and the geometric coefficients of fullsky_forms.h (the reference's
    // polynomials in cos theta with their (chi/s)^n cancellations done exactly)
    endpoint e1, e2;
    ce_fullsky_geom g;
    if (!a.flatsky) {
        e1 = background_at(a, chi1, jbar, chib, dchi);
        e2 = background_at(a, chi2, jbar, chib, dchi);
        g = ce_fullsky_geometry(chi_mean, sep, mu);
    } else {
        e1.z = e1.f = e1.H = e1.G1 = e1.G2 = e1.D1 = e1.b1 = e1.b2 = 0.0;
        e2 = e1;
    }
    const double f1 = e1.f, f2 = e2.f, H1 = e1.H, H2 = e2.H, b1 = e1.b1, b2 = e2.b2;
    const double G1 = e1.G1, G2 = e2.G2;
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
    const double G = a.flatsky ? Dmean * Dmean : e1.D1 * e2.D1;
    for (int s = 0; s < NSTD; ++s) c[s] *= G;
}

// one wavefront per (z, r); blockDim = 256 (four points).  The lanes take 64 mu nodes
// at a time; after each batch every (multipole, slot) sum is reduced across the
// wavefront and added into the wavefront's LDS accumulators, so no per-lane partial
// sums stay live over the node loop and the kernel fits 128 VGPRs (4 waves per SIMD,
// every (z, r) wavefront of the 8 x 512 grid resident at once) without spills.
// NL = the number of multipoles kept (>= a.nl), a template argument so that absent
// multipoles cost nothing
template <int NL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_mu_weights(mu_args a)
{
    __shared__ double sacc[4][CE_MU_MAXL * NSTD];
    const int lane = threadIdx.x & 63;
    const int point = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int npoint = a.nz * a.nsep;
    if (point >= npoint) return;   // whole wavefronts exit together
    double *acc = sacc[threadIdx.x >> 6];
    if (lane < NL * NSTD) acc[lane] = 0.0;
    const int iz = point / a.nsep, ir = point % a.nsep;
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
        double c[NSTD];
        double mu = 0.0, w = 0.0;
        if (q < a.nmu) {
            mu = a.mu[q];
            w = a.wmu[q];
            local_coefficients(a, zm, mu, sep, chi_mean, fmean, Hmean, bm1, bm2, Gm1, Gm2, Dmean, jbar, chib, dchi, c);
        } else {
#pragma unroll
            for (int s = 0; s < NSTD; ++s) c[s] = 0.0;
        }
#pragma unroll
        for (int il = 0; il < NL; ++il) {
            if (il >= a.nl) break;
            const double pl = w * legendre_p(a.ls[il], mu);
#pragma unroll
            for (int s = 0; s < NSTD; ++s) {
                double v = pl * c[s];
                for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
                if (lane == 0) acc[il * NSTD + s] += v;
            }
        }
    }
    // lane 0 wrote the sums; the other lanes read them after the wavefront's LDS writes
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0): the LDS stores have completed
    if (lane < NL * NSTD) {
        const int il = lane / NSTD, s = lane % NSTD;
        if (il < a.nl) {
            const double v = acc[lane];
            for (int t = 0; t < a.ntab; ++t)
                if (a.slot[t] == s) a.A[(((int64_t)iz * a.nl + il) * a.ntab + t) * a.nsep + ir] = (2 * a.ls[il] + 1) * 0.5 * v;
        }
    }
}

}  // namespace

extern "C" int ce_launch_mu_weights(void *stream, int nbg, const double *bg_z, const double *bg_chi,
                                    const double *bg_D1, const double *bg_f, const double *bg_cH, const double *bg_G1,
                                    const double *bg_G2, int nb, const double *bias_z, const double *bias1,
                                    const double *bias2, unsigned terms, int flatsky, int nz,
                                    const double *zmean, int nl, const int *ls, int ntab, const int *slot, int nsep,
                                    const double *sep, int nmu, const double *mu, const double *wmu, double *A,
                                    double *coef)
{
    if (nl > CE_MU_MAXL || ntab > CE_MAX_TABLES) return (int)hipErrorInvalidValue;
    mu_args a;
    coef_args ca;
    {
        const double *xs[CE_MU_NCOEF] = {bg_chi, bg_z, bg_z, bg_z, bg_z, bg_z, bias_z, bias_z};
        const double *ys[CE_MU_NCOEF] = {bg_z, bg_f, bg_cH, bg_G1, bg_G2, bg_D1, bias1, bias2};
        double *p = coef;
        int nmax = 0;
        for (int t = 0; t < CE_MU_NCOEF; ++t) {
            const int n = t < 6 ? nbg : nb;
            ca.x[t] = xs[t]; ca.y[t] = ys[t]; ca.n[t] = n;
            ca.b[t] = p; ca.c[t] = p + n; ca.d[t] = p + 2 * n;
            a.cb[t] = ca.b[t]; a.cc[t] = ca.c[t]; a.cd_[t] = ca.d[t];
            p += 3 * n;
            if (n > nmax) nmax = n;
        }
        hipLaunchKernelGGL(k_akima_coef, dim3((nmax + 255) / 256, CE_MU_NCOEF), dim3(256), 0, (hipStream_t)stream, ca);
    }
    a.nbg = nbg; a.bg_z = bg_z; a.bg_chi = bg_chi; a.bg_D1 = bg_D1; a.bg_f = bg_f; a.bg_cH = bg_cH;
    a.bg_G1 = bg_G1; a.bg_G2 = bg_G2;
    a.nb = nb; a.bias_z = bias_z; a.bias1 = bias1; a.bias2 = bias2;
    a.terms = terms; a.flatsky = flatsky;
    a.nz = nz; a.zmean = zmean; a.nl = nl;
    for (int i = 0; i < nl; ++i) a.ls[i] = ls[i];
    a.ntab = ntab;
    for (int t = 0; t < ntab; ++t) a.slot[t] = slot[t];
    a.nsep = nsep; a.sep = sep; a.nmu = nmu; a.mu = mu; a.wmu = wmu; a.A = A;
    const int npoint = nz * nsep;
    const int blocks = (npoint + 3) / 4;
    if (nl <= 1) hipLaunchKernelGGL(k_mu_weights<1>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, a);
    else if (nl == 2) hipLaunchKernelGGL(k_mu_weights<2>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, a);
    else if (nl == 3) hipLaunchKernelGGL(k_mu_weights<3>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, a);
    else hipLaunchKernelGGL(k_mu_weights<CE_MU_MAXL>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, a);
    return (int)hipGetLastError();
}
