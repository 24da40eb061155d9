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

#define CE_MU_MAXL 4

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
};

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
    e.z = ce::akima_at(a.bg_chi, a.bg_z, n, jc, chi);
    const int k = ce::interval_from(a.bg_z, n, e.z, ce::interval_guess(a.bg_z, n, e.z));
    e.f = ce::akima_at(a.bg_z, a.bg_f, n, k, e.z);
    e.H = ce::akima_at(a.bg_z, a.bg_cH, n, k, e.z);
    e.G1 = ce::akima_at(a.bg_z, a.bg_G1, n, k, e.z);
    e.G2 = ce::akima_at(a.bg_z, a.bg_G2, n, k, e.z);
    e.D1 = ce::akima_at(a.bg_z, a.bg_D1, n, k, e.z);
    if (!(e.z >= a.bias_z[0] && e.z <= a.bias_z[a.nb - 1])) {
        e.b1 = e.b2 = __builtin_nan("");
    } else {
        const int kb = ce::interval_from(a.bias_z, a.nb, e.z, ce::interval_guess(a.bias_z, a.nb, e.z));
        e.b1 = ce::akima_at(a.bias_z, a.bias1, a.nb, kb, e.z);
        e.b2 = ce::akima_at(a.bias_z, a.bias2, a.nb, kb, e.z);
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
    const double costheta = (2. * chi_mean * chi_mean - sep * sep + mu * mu * sep * sep / 2.) /
                            (2. * chi_mean * chi_mean - mu * mu * sep * sep / 2.);
    const double ct2 = costheta * costheta;
    // the flat-sky forms use the mean's background only
    endpoint e1, e2;
    if (!a.flatsky) {
        e1 = background_at(a, chi1, jbar, chib, dchi);
        e2 = background_at(a, chi2, jbar, chib, dchi);
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
            c[S00] += f1 * f2 * (1 + 2 * ct2) / 15;
            c[S02] += -f1 * f2 / 21. * ((1 + 11. * ct2) + 18 * costheta * (ct2 - 1) * chi1 * chi2 / sep / sep);
            c[S04] += f1 * f2 *
                      (4 * (3 * ct2 - 1) * (pow(chi1, 4) + pow(chi2, 4)) / 35. / pow(sep, 4) +
                       chi1 * chi2 * (3 + ct2) * (3 * (3 + ct2) * chi1 * chi2 - 8 * (chi1 * chi1 + chi2 * chi2) * costheta) /
                           35. / pow(sep, 4));
        }
    }
    if (a.terms & COFFE_TERM_D1_D1) {   // functions.c:363-573
        if (a.flatsky) {
            c[S20] += Hmean * Hmean * fmean * fmean * Gm1 * Gm2 / 3.;
            c[S22] += 2 * Hmean * Hmean * fmean * fmean * Gm1 * Gm2 * pow(sep, 2) * P2 / 3.;
        } else {
            c[S20] += H1 * H2 * f1 * f2 * G1 * G2 * costheta / 3.;
            c[S22] += -H1 * H2 * f1 * f2 * G1 * G2 *
                      ((chi2 - chi1 * costheta) * (chi1 - chi2 * costheta) + pow(sep, 2) * costheta / 3.);
        }
    }
    if (a.terms & COFFE_TERM_DEN_RSD) {   // functions.c:670-771
        if (a.flatsky) {
            c[S00] += bm1 * fmean / 3. + bm2 * fmean / 3.;
            c[S02] += -(2 * bm1 * fmean / 3. + 2 * bm2 * fmean / 3.) * P2;
        } else {
            c[S00] += b1 * f2 / 3. + b2 * f1 / 3.;
            c[S02] += -(b1 * f2 * (2. / 3. - (1. - ct2) * pow(chi1 / sep, 2)) +
                        b2 * f1 * (2. / 3. - (1. - ct2) * pow(chi2 / sep, 2)));
        }
    }
    if (a.terms & COFFE_TERM_DEN_D1) {   // functions.c:772-855
        if (a.flatsky) {
            c[S11] += (bm1 * fmean * Hmean * Gm2 * mu + bm2 * fmean * Hmean * Gm1 * (-mu)) * sep;
        } else {
            c[S11] += -(b1 * f2 * H2 * G2 * (chi1 * costheta - chi2) + b2 * f1 * H1 * G1 * (chi2 * costheta - chi1));
        }
    }
    if (a.terms & COFFE_TERM_RSD_D1) {   // functions.c:904-1143
        if (a.flatsky) {
            const double P3p = legendre_p(3, mu), P3m = legendre_p(3, -mu);
            c[S11] += -3. / 5. * (fmean * fmean * Hmean * Gm2 * (-mu) + fmean * fmean * Hmean * Gm1 * mu) * sep;
            c[S13] += (fmean * fmean * Hmean * Gm2 * 2. / 5. * P3m + fmean * fmean * Hmean * Gm1 * 2. / 5. * P3p) * sep;
        } else {
            c[S11] += f1 * f2 * H2 * G2 * ((1. + 2 * ct2) * chi2 - 3 * chi1 * costheta) / 5. +
                      f2 * f1 * H1 * G1 * ((1. + 2 * ct2) * chi1 - 3 * chi2 * costheta) / 5.;
            c[S13] += (f1 * f2 * H2 * G2 *
                           ((1. - 3 * ct2) * pow(chi2, 3) + costheta * (5. + ct2) * pow(chi2, 2) * chi1 -
                            2 * (2. + ct2) * chi2 * pow(chi1, 2) + 2 * pow(chi1, 3) * costheta) /
                           5 +
                       f2 * f1 * H1 * G1 *
                           ((1. - 3 * ct2) * pow(chi1, 3) + costheta * (5. + ct2) * pow(chi1, 2) * chi2 -
                            2 * (2. + ct2) * chi1 * pow(chi2, 2) + 2 * pow(chi2, 3) * costheta) /
                           5) /
                      pow(sep, 2);
        }
    }
    const double G = a.flatsky ? Dmean * Dmean : e1.D1 * e2.D1;
    for (int s = 0; s < NSTD; ++s) c[s] *= G;
}

// one wavefront per (z, r); blockDim = 256 (four points)
__global__ void __launch_bounds__(256) k_mu_weights(mu_args a)
{
    const int lane = threadIdx.x & 63;
    const int point = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int npoint = a.nz * a.nsep;
    if (point >= npoint) return;   // whole wavefronts exit together
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
    // partial sums per (multipole, standard table slot); fixed bounds keep them in registers
    double acc[CE_MU_MAXL][NSTD];
#pragma unroll
    for (int l = 0; l < CE_MU_MAXL; ++l)
#pragma unroll
        for (int s = 0; s < NSTD; ++s) acc[l][s] = 0.0;
    for (int q = lane; q < a.nmu; q += 64) {
        const double mu = a.mu[q], w = a.wmu[q];
        double c[NSTD];
        local_coefficients(a, zm, mu, sep, chi_mean, fmean, Hmean, bm1, bm2, Gm1, Gm2, Dmean, jbar, chib, dchi, c);
#pragma unroll
        for (int il = 0; il < CE_MU_MAXL; ++il) {
            if (il < a.nl) {
                const double pl = w * legendre_p(a.ls[il], mu);
#pragma unroll
                for (int s = 0; s < NSTD; ++s) acc[il][s] += pl * c[s];
            }
        }
    }
#pragma unroll
    for (int il = 0; il < CE_MU_MAXL; ++il) {
        if (il >= a.nl) break;
#pragma unroll
        for (int s = 0; s < NSTD; ++s) {
            double v = acc[il][s];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) {
                for (int t = 0; t < a.ntab; ++t)
                    if (a.slot[t] == s)
                        a.A[(((int64_t)iz * a.nl + il) * a.ntab + t) * a.nsep + ir] = (2 * a.ls[il] + 1) * 0.5 * v;
            }
        }
    }
}

}  // namespace

extern "C" int ce_launch_mu_weights(void *stream, int nbg, const double *bg_z, const double *bg_chi,
                                    const double *bg_D1, const double *bg_f, const double *bg_cH, const double *bg_G1,
                                    const double *bg_G2, int nb, const double *bias_z, const double *bias1,
                                    const double *bias2, unsigned terms, int flatsky, int nz,
                                    const double *zmean, int nl, const int *ls, int ntab, const int *slot, int nsep,
                                    const double *sep, int nmu, const double *mu, const double *wmu, double *A)
{
    if (nl > CE_MU_MAXL || ntab > CE_MAX_TABLES) return (int)hipErrorInvalidValue;
    mu_args a;
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
    hipLaunchKernelGGL(k_mu_weights, dim3(blocks), dim3(256), 0, (hipStream_t)stream, a);
    return (int)hipGetLastError();
}
