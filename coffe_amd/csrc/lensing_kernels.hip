// Flat-sky lensing multipoles on the GPU (gfx950, FP64).
//
// The two lensing terms the reference evaluates without Monte Carlo when the flat-sky
// switches are on (signal.c:612-628 and 704-736):
//   - lensing-lensing, even l (functions.c:4020-4133): 2 pi^2 times
//       int_0^1 dx (2l+1) l! / (2^l ((l/2)!)^2) 9 Omega_m^2 (2 - 5 s1)(2 - 5 s2) chi^3 / (8 pi)
//                  * r I_l^1(x r) D1(z(x chi))^2 (1 + z(x chi))^2 (x (1 - x))^2 (H0 h)^4
//     with chi = chi(z_mean), r I_l^1 the (n, l) = (1, l) table (r^(n-l) I_l^n stored for
//     n >= l, times x r here for l > 0), z(lambda) the Akima spline of z against chi
//     and D1 that of D1 against z;
//   - density-lensing, even l (functions.c:4136-4257): 2 pi^2 times
//       sum_{k <= l/2} (-1)^k / 2^k C(l, k) C(2l - 2k, l) (l/2 - k)! I_k(r)
//       * -3 Omega_m / (8 pi) (2l+1) D1^2 (1 + z) r ((2 - 5 s1) b2 + (2 - 5 s2) b1)
//       * 2^(1 - l/2) / sqrt(2 pi) (H0 h)^2,
//     with I_k the half-integer table (n, l) = ((l - 2k + 3)/2, (l - 2k + 1)/2).
// Odd multipoles of both are zero (the reference's construction).
//
// One wavefront per (z, l, r): the lanes take the quadrature nodes of the x integral
// (a composite Gauss-Legendre rule on [0, 1]) and reduce by lane shuffles; lane 0
// adds the closed-form term.  Every table lookup is an Akima interpolation of raw
// nodes (ce::akima_general, GSL's non-rounded Akima), as the reference's splines.

#include <hip/hip_runtime.h>

#include "ce_device.h"
#include "ce_internal.h"

namespace {

struct flatlens_args {
    int mode;                            // 0: multipoles over ls; 1: correlation function over mus
    const double *mus;                   // [nl] in mode 1
    int nz, nl, nsep, nb, ngl, want_ll, want_dl;
    const double *zpar;                  // [nz][7]: z, chi, D1, s1, s2, b1, b2 at z_mean
    const int *ls;                       // [nl]
    const double *sep;                   // [nsep]
    const double *bg_z, *bg_chi, *bg_D1; // [nb] background samples (chi increasing)
    const double *tab_x, *tab_y;         // the tables, concatenated
    const int64_t *ll_off;               // [nl] offset of the (1, l) table
    const int *ll_len;                   // [nl] its length (0: none)
    const int64_t *dl_off;               // [nl][CE_FL_MAXK] offsets of the half-integer tables
    const int *dl_len;                   // [nl][CE_FL_MAXK] their lengths (0: none)
    const double *glx, *glw;             // [ngl] rule on [0, 1]
    double Omega_m, H0h;
    double *out;                         // [nz][nl][nsep]
};

__device__ inline double factorial(int n)
{
    double f = 1.0;
    for (int i = 2; i <= n; ++i) f *= i;
    return f;
}

__device__ inline double binomial(int n, int k)
{
    double c = 1.0;
    for (int i = 1; i <= k; ++i) c = c * (n - k + i) / i;
    return c;
}

// Mode 1, the correlation function at (z, mu, r) (functions.c:3515-3558 and 1563-1622,
// the branches for flatsky_nonlocal / flatsky_local_nonlocal outside MULTIPOLES): with
// T the (1/2, -1/2) half-integer table (r I(r)) times sqrt(2/pi) 2 pi^2 (integrals.c:956-
// 962) and s = sqrt(1 - mu^2),
//   lensing-lensing = int_0^1 dx 9 Omega_m^2 (2 - 5 s1)(2 - 5 s2) chi^3 / (8 pi)
//                     T(x r s) D1(z(x chi))^2 (1 + z(x chi))^2 (x (1 - x))^2 (H0 h)^4
//     (the reference's second line-of-sight variable does not enter: its integral is 1),
//   density-lensing = (H0 h)^2 (-3 Omega_m / (8 pi)) D1^2 (1 + z) |mu| r
//                     ((2 - 5 s1) b2 + (2 - 5 s2) b1) T(r s).
__global__ void __launch_bounds__(64) k_flat_lensing_cf(flatlens_args a)
{
    const int idx = blockIdx.x;   // (z, mu, r), r fastest
    const int ir = idx % a.nsep;
    const int im = (idx / a.nsep) % a.nl;
    const int iz = idx / (a.nsep * a.nl);
    const int lane = threadIdx.x;
    const double mu = a.mus[im];
    const double r = a.sep[ir];
    const double sm = sqrt(fmax(1.0 - mu * mu, 0.0));
    const double *zp = a.zpar + 7 * iz;
    const double z = zp[0], chi = zp[1], D1 = zp[2], s1 = zp[3], s2 = zp[4], b1 = zp[5], b2 = zp[6];
    const double pi = 3.14159265358979323846;
    const double *tx = a.tab_x + a.ll_off[0], *ty = a.tab_y + a.ll_off[0];
    const int tn = a.ll_len[0];
    const double tscale = sqrt(2.0 / pi) * 2.0 * pi * pi;
    double acc = 0.0;
    if (a.want_ll) {
        for (int q = lane; q < a.ngl; q += 64) {
            const double x = a.glx[q];
            const double zl = ce::akima_general(a.bg_chi, a.bg_z, a.nb, chi * x);
            const double Dl = ce::akima_general(a.bg_z, a.bg_D1, a.nb, zl);
            const double t = tscale * ce::akima_general(tx, ty, tn, x * r * sm);
            const double xx = x * (1.0 - x);
            acc += a.glw[q] * t * Dl * Dl * (1.0 + zl) * (1.0 + zl) * xx * xx;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
        const double h2 = a.H0h * a.H0h;
        acc *= 9.0 * a.Omega_m * a.Omega_m * (2.0 - 5.0 * s1) * (2.0 - 5.0 * s2) * chi * chi * chi / 8.0 / pi * h2 * h2;
    }
    if (lane != 0) return;
    if (a.want_dl)
        acc += a.H0h * a.H0h * (-3.0 * a.Omega_m / pi / 8.0 * D1 * D1 * (1.0 + z) * fabs(mu) * r *
                                ((2.0 - 5.0 * s1) * b2 + (2.0 - 5.0 * s2) * b1) *
                                (tscale * ce::akima_general(tx, ty, tn, r * sm)));
    a.out[idx] = acc;
}

__global__ void __launch_bounds__(64) k_flat_lensing(flatlens_args a)
{
    const int idx = blockIdx.x;   // (z, l, r), r fastest
    const int ir = idx % a.nsep;
    const int il = (idx / a.nsep) % a.nl;
    const int iz = idx / (a.nsep * a.nl);
    const int lane = threadIdx.x;
    const int l = a.ls[il];
    const double r = a.sep[ir];
    const double *zp = a.zpar + 7 * iz;
    const double z = zp[0], chi = zp[1], D1 = zp[2], s1 = zp[3], s2 = zp[4], b1 = zp[5], b2 = zp[6];
    const double pi = 3.14159265358979323846;
    double acc = 0.0;
    const bool even = (l % 2) == 0;
    if (a.want_ll && even && a.ll_len[il] > 0) {
        const double *tx = a.tab_x + a.ll_off[il], *ty = a.tab_y + a.ll_off[il];
        const int tn = a.ll_len[il];
        for (int q = lane; q < a.ngl; q += 64) {
            const double x = a.glx[q];
            const double zl = ce::akima_general(a.bg_chi, a.bg_z, a.nb, chi * x);
            const double Dl = ce::akima_general(a.bg_z, a.bg_D1, a.nb, zl);
            double t = ce::akima_general(tx, ty, tn, x * r);
            if (l > 0) t *= x * r;
            const double xx = x * (1.0 - x);
            acc += a.glw[q] * t * Dl * Dl * (1.0 + zl) * (1.0 + zl) * xx * xx;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
        const double h2 = a.H0h * a.H0h;
        acc *= (2 * l + 1) * factorial(l) / ldexp(1.0, l) / (factorial(l / 2) * factorial(l / 2)) * 9.0 *
               a.Omega_m * a.Omega_m * (2.0 - 5.0 * s1) * (2.0 - 5.0 * s2) * chi * chi * chi / 8.0 / pi * h2 * h2 *
               (2.0 * pi * pi);
    }
    if (lane != 0) return;
    if (a.want_dl && even) {
        double s = 0.0;
        bool ok = true;
        for (int k = 0; k <= l / 2 && k < CE_FL_MAXK; ++k) {
            const int t = il * CE_FL_MAXK + k;
            if (a.dl_len[t] <= 0) { ok = false; break; }
            const double v = ce::akima_general(a.tab_x + a.dl_off[t], a.tab_y + a.dl_off[t], a.dl_len[t], r);
            s += ((k & 1) ? -1.0 : 1.0) / ldexp(1.0, k) * binomial(l, k) * binomial(2 * l - 2 * k, l) *
                 factorial(l / 2 - k) * v;
        }
        if (!ok || l / 2 >= CE_FL_MAXK) s = __builtin_nan("");
        s *= -3.0 * a.Omega_m / pi / 8.0 * (2 * l + 1) * D1 * D1 * (1.0 + z) * r *
             ((2.0 - 5.0 * s1) * b2 + (2.0 - 5.0 * s2) * b1) * ldexp(1.0, 1 - l / 2);
        acc += 2.0 * pi * pi * s / sqrt(2.0 * pi) * a.H0h * a.H0h;
    }
    a.out[idx] = acc;
}

}  // namespace

extern "C" int ce_launch_flat_lensing(void *stream, int mode, const double *d_mus, int nz, int nl, int nsep, int nb,
                                      int ngl, int want_ll, int want_dl, const double *d_zpar, const int *d_ls,
                                      const double *d_sep,
                                      const double *d_bg_z, const double *d_bg_chi, const double *d_bg_D1,
                                      const double *d_tab_x, const double *d_tab_y, const int64_t *d_ll_off,
                                      const int *d_ll_len, const int64_t *d_dl_off, const int *d_dl_len,
                                      const double *d_glx, const double *d_glw, double Omega_m, double H0h,
                                      double *d_out)
{
    flatlens_args a;
    a.mode = mode;
    a.mus = d_mus;
    a.nz = nz;
    a.nl = nl;
    a.nsep = nsep;
    a.nb = nb;
    a.ngl = ngl;
    a.want_ll = want_ll;
    a.want_dl = want_dl;
    a.zpar = d_zpar;
    a.ls = d_ls;
    a.sep = d_sep;
    a.bg_z = d_bg_z;
    a.bg_chi = d_bg_chi;
    a.bg_D1 = d_bg_D1;
    a.tab_x = d_tab_x;
    a.tab_y = d_tab_y;
    a.ll_off = d_ll_off;
    a.ll_len = d_ll_len;
    a.dl_off = d_dl_off;
    a.dl_len = d_dl_len;
    a.glx = d_glx;
    a.glw = d_glw;
    a.Omega_m = Omega_m;
    a.H0h = H0h;
    a.out = d_out;
    const int64_t blocks = (int64_t)nz * nl * nsep;
    if (blocks <= 0 || blocks > 0x7fffffff) return (int)hipErrorInvalidValue;
    if (mode == 1)
        hipLaunchKernelGGL(k_flat_lensing_cf, dim3((unsigned)blocks), dim3(64), 0, (hipStream_t)stream, a);
    else
        hipLaunchKernelGGL(k_flat_lensing, dim3((unsigned)blocks), dim3(64), 0, (hipStream_t)stream, a);
    return (int)hipGetLastError();
}
