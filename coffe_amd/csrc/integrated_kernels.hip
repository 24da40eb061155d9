// Integrated (lensing and potential) terms of the correlation function, full sky, on
// the GPU (gfx950, FP64).
//
// The reference evaluates them with CUBA's Cuhre or GSL Monte Carlo over [0, 1]^2
// (double-integrated, functions.c:3460-4017, signal.c:735-741) and with QAG or a
// double-exponential rule over [0, 1] (single-integrated, functions.c:1486-3458,
// signal.c:578-600).  Here both are deterministic product rules, one workgroup per
// (z_mean, mu, separation):
//   - double: x1 a composite Gauss-Legendre rule on [0, 1] (cut where lambda1 cos(theta)
//     = chi2), and for each x1 the substitution lambda2 = lambda1 c + s sinh(u),
//     s = lambda1 sin(theta), under which r = s cosh(u) and the ridge of closest
//     approach along lambda2 is smooth in u; u a composite Gauss-Legendre rule;
//   - single: the two mixed separations are closest at x = chi1 c / chi2 and chi2 c /
//     chi1; [0, 1] is cut there (and halfway between), each piece mapped by x = centre
//     +- s sinh(u) about its nearer centre (s the ridge width), u Gauss-Legendre.
// Every lookup is an Akima interpolation of raw nodes (ce::akima_general) and the
// renormalisation grid a GSL-style bicubic, as the reference's splines.

#include <hip/hip_runtime.h>

#include "ce_device.h"
#include "ce_internal.h"

namespace {

enum { S00, S02, S04, S11, S13, S20, S22, S31, S40, NSLOT };
// background curves with precomputed Akima coefficients: D1, a, f, conformal H against z,
// and z against chi
enum { C_D1, C_A, C_F, C_CH, C_ZOF, NBGC };

constexpr unsigned F_DEN = 1u << 0, F_RSD = 1u << 1, F_D1 = 1u << 2, F_D2 = 1u << 3, F_G1 = 1u << 4,
                   F_G2 = 1u << 5, F_G3 = 1u << 6, F_LEN = 1u << 7, F_G4 = 1u << 8, F_G5 = 1u << 9,
                   F_ONLY_CROSS = 1u << 10, F_FLAT_LL = 1u << 11, F_FLAT_DL = 1u << 12;

struct intg_args {
    int nbg;
    const double *bg_z, *bg_chi, *bg_D1, *bg_a, *bg_f, *bg_cH, *bg_G1, *bg_G2;
    int nb;
    const double *bias_z, *b1, *b2;
    int nmb;
    const double *mb_z, *s1, *s2, *fe1, *fe2;
    const double *tab_x, *tab_y;
    int tab_off[NSLOT], tab_len[NSLOT];
    int tab_f0[NSLOT];                 // first node of the log-uniform part (-1: bisection)
    int tab_ref;                       // a slot whose nodes every present standard table shares (-1: none)
    double tab_lx0[NSLOT], tab_lsc[NSLOT];   // its log(x[f0]) and nodes per unit of log x
    int nguess;                        // chi -> interval guesses on [0, chi_max], nguess + 1 of them
    const double *chi_guess;
    // Akima value records and node pairs (below): the tables' in the layout of tab_y,
    // the background curves' [NBGC][nbg]
    const double *tab_val, *tab_pair, *bgc_val, *bgc_pair;
    int nren0;
    const double *ren0_x, *ren0_y;
    int nren;
    const double *ren_chi;
    const double *ren_rec;   // per grid node (x fastest): {value, d/dx, d/dy, d2/dxdy}
    double Om, H0h;
    unsigned flags;
    int nz, nmu, nsep;
    const double *zmean, *mu, *sep;
    int p1, n1, pu, nu, ps, ns;
    const double *g1x, *g1w, *gux, *guw, *gsx, *gsw;   // Gauss-Legendre rules on [-1, 1]
    double *out;                                       // [nz][nmu][nsep][2]: single, double
};

__device__ inline double ak(const double *x, const double *y, int n, double v) { return ce::akima_general(x, y, n, v); }

// An Akima lookup reads two arrays: node pairs {x_j, x_(j+1)} (16 bytes, one load: both
// ends of interval j, which also test a guessed interval), shared by the curves on one
// grid, and each curve's value records {y_j, b_j, c_j, d_j} (32 bytes, two loads).  The
// lookups of this kernel gather at random per lane, so the L1 cache's throughput in
// cache lines, not arithmetic, bounds them: separate node and coefficient arrays cost
// five scattered lines per lookup, this layout three, and two for each further curve
// or table on the same interval.
constexpr int VAL = CE_INTG_VAL, PAIR = CE_INTG_PAIR;
struct nodes {
    double x0, x1;
};

__device__ __forceinline__ nodes pair_at(const double *__restrict__ p, int j)
{
    const double2 q = *reinterpret_cast<const double2 *>(p + (size_t)PAIR * j);
    return nodes{q.x, q.y};
}

// whether interval j (0 <= j <= n - 2), whose nodes g holds, is the one
// ce::interval_from returns from the guess j at once (false for a NaN v, whose callers
// then search and get j back)
__device__ __forceinline__ bool pair_hit(const nodes &g, int j, int n, double v)
{
    return g.x0 <= v && (j == n - 2 || v < g.x1);
}

// y(xq) on the segment of interval j with lower node x0 (the value akima_at gives there)
__device__ __forceinline__ double val_at(const double *__restrict__ r, int j, double x0, double xq)
{
    const double2 p = *reinterpret_cast<const double2 *>(r + (size_t)VAL * j),
                  q = *reinterpret_cast<const double2 *>(r + (size_t)VAL * j + 2);
    const double dx = xq - x0;
    return p.x + dx * (p.y + dx * (q.x + q.y * dx));
}

// The lookups: Akima of raw nodes as ce::akima_general, with the interval found from a
// guess (the background's redshift grid is uniform; chi(z) through a table of interval
// guesses; the integral tables' FFT nodes are log-uniform) and a short walk, falling
// back to bisection: the same interval, so the same value, in a few dependent loads
// where bisection over 10^4 nodes takes 14.
struct lk {
    const intg_args &a;
    __device__ lk(const intg_args &a_) : a(a_) {}
    __device__ double bg(const double *col, double z) const
    {
        const double *x = a.bg_z;
        const int n = a.nbg;
        if (!(z >= x[0] && z <= x[n - 1])) return __builtin_nan("");
        return ce::akima_at(x, col, n, ce::interval_from(x, n, z, ce::interval_guess(x, n, z)), z);
    }
    static __device__ int below_last(int g, int n) { return g > n - 2 ? n - 2 : g; }
    // the first guess of z's interval on the background's redshift grid (-1 outside it)
    __device__ int bg_guess(double z) const
    {
        const double *x = a.bg_z;
        const int n = a.nbg;
        if (!(z >= x[0] && z <= x[n - 1])) return -1;
        return below_last(ce::interval_guess(x, n, z), n);
    }
    // the interval of z on the redshift grid (-1 outside it), which the curves read at
    // one z share
    __device__ int bg_interval(double z) const
    {
        const int g = bg_guess(z);
        return g < 0 ? -1 : ce::interval_from(a.bg_z, a.nbg, z, g);
    }
    // curve k's value records, and the node pairs of its grid (bg_z for D1, a, f and
    // conformal H; bg_chi for z(chi))
    __device__ const double *bgv(int k) const { return a.bgc_val + (size_t)VAL * k * a.nbg; }
    __device__ const double *bgp(int k) const
    {
        return a.bgc_pair + (size_t)PAIR * (k == C_ZOF ? C_ZOF : C_D1) * a.nbg;
    }
    // D1, a, f or conformal H (curve k) in interval j (NaN for j < 0)
    __device__ double bgcj(int k, int j, double z) const
    {
        return j < 0 ? __builtin_nan("") : val_at(bgv(k), j, pair_at(bgp(k), j).x0, z);
    }
    // curve k at z, the interval found from the guess; *jz (if given) receives it
    __device__ double bgc(int k, double z, int *jz = nullptr) const
    {
        int j = bg_guess(z);
        double v = __builtin_nan("");
        if (j >= 0) {
            const nodes g = pair_at(bgp(k), j);
            v = val_at(bgv(k), j, g.x0, z);
            if (!pair_hit(g, j, a.nbg, z)) {
                j = ce::interval_from(a.bg_z, a.nbg, z, j);
                v = bgcj(k, j, z);
            }
        }
        if (jz) *jz = j;
        return v;
    }
    __device__ double z_of(double chi) const
    {
        const double *x = a.bg_chi;
        const int n = a.nbg;
        if (!(chi >= x[0] && chi <= x[n - 1])) return __builtin_nan("");
        const int k = ce::clamp_index(chi / x[n - 1] * a.nguess, a.nguess + 1);
        const int j = below_last(max((int)a.chi_guess[k], 0), n);
        const nodes g = pair_at(bgp(C_ZOF), j);
        const double v = val_at(bgv(C_ZOF), j, g.x0, chi);
        if (pair_hit(g, j, n, chi)) return v;
        return bgcj(C_ZOF, ce::interval_from(x, n, chi, j), chi);
    }
    // a first interval of r on table s's nodes (-1: r outside them): by bisection below
    // the log-uniform part (exact), from the log above it (a guess)
    __device__ int tab_guess(int s, double r) const
    {
        const double *x = a.tab_x + a.tab_off[s];
        const int n = a.tab_len[s], f0 = a.tab_f0[s];
        if (!(r >= x[0] && r <= x[n - 1])) return -1;
        if (f0 < 0) return ce::interval_bisect(x, 0, n - 1, r);
        if (r < x[f0]) return ce::interval_bisect(x, 0, f0, r);
        return below_last(f0 + ce::clamp_index((log(r) - a.tab_lx0[s]) * a.tab_lsc[s], n - f0), n);
    }
    __device__ const double *tabv(int s) const { return a.tab_val + (size_t)VAL * a.tab_off[s]; }
    __device__ const double *tabp(int s) const { return a.tab_pair + (size_t)PAIR * a.tab_off[s]; }
    // interval j of table s settled from its guess g (-1 stays -1); *x0 its lower node
    __device__ int tab_settle(int s, int g, double r, double *x0) const
    {
        if (g < 0) return g;
        nodes nd = pair_at(tabp(s), g);
        if (!pair_hit(nd, g, a.tab_len[s], r)) {
            g = ce::interval_from(a.tab_x + a.tab_off[s], a.tab_len[s], r, g);
            nd = pair_at(tabp(s), g);
        }
        *x0 = nd.x0;
        return g;
    }
    __device__ double tab(int s, double r) const
    {
        const int g = tab_guess(s, r);
        if (g < 0) return __builtin_nan("");
        const nodes nd = pair_at(tabp(s), g);
        const double v = val_at(tabv(s), g, nd.x0, r);
        if (pair_hit(nd, g, a.tab_len[s], r)) return v;
        const int j = ce::interval_from(a.tab_x + a.tab_off[s], a.tab_len[s], r, g);
        return val_at(tabv(s), j, pair_at(tabp(s), j).x0, r);
    }
    // the interval of r on the standard tables' shared nodes (-1: none shared, or r
    // outside them) and its lower node; a standard table's value there
    __device__ int tab_interval(double r, double *x0) const
    {
        return a.tab_ref < 0 ? -1 : tab_settle(a.tab_ref, tab_guess(a.tab_ref, r), r, x0);
    }
    __device__ double tabj(int s, int j, double x0, double r) const
    {
        return j < 0 ? tab(s, r) : val_at(tabv(s), j, x0, r);
    }
    // GSL interp2d bicubic of the renormalisation grid at (x, y) (node derivatives of
    // natural cubic splines, Hermite patches; common.c:712-719)
    __device__ double bicubic(double x, double y) const
    {
        const int n = a.nren;
        const double *c = a.ren_chi;
        if (!(x >= c[0] && x <= c[n - 1] && y >= c[0] && y <= c[n - 1])) return __builtin_nan("");
        const int i = ce::interval_from(c, n, x, ce::interval_guess(c, n, x));
        const int j = ce::interval_from(c, n, y, ce::interval_guess(c, n, y));
        const double dx = c[i + 1] - c[i], dy = c[j + 1] - c[j];
        const double t = (x - c[i]) / dx, u = (y - c[j]) / dy;
        const double ht[4] = {(1 + 2 * t) * (1 - t) * (1 - t), t * t * (3 - 2 * t), t * (1 - t) * (1 - t),
                              t * t * (t - 1)};
        const double hu[4] = {(1 + 2 * u) * (1 - u) * (1 - u), u * u * (3 - 2 * u), u * (1 - u) * (1 - u),
                              u * u * (u - 1)};
        double r = 0.0;
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const double *g = a.ren_rec + 4 * ((size_t)(j + q) * n + i + p);
                const double2 vx = *reinterpret_cast<const double2 *>(g), yxy = *reinterpret_cast<const double2 *>(g + 2);
                r += vx.x * ht[p] * hu[q] + vx.y * dx * ht[2 + p] * hu[q] + yxy.x * dy * ht[p] * hu[2 + q] +
                     yxy.y * dx * dy * ht[2 + p] * hu[2 + q];
            }
        return r;
    }
    // I_0^4 r^4 at r minus the grid at (x, y); the r = 0 renormalisation at `at0` where
    // `small` holds (functions.c:1524-1553, 3494-3512)
    __device__ double ren(double r2, double x, double y, bool small, double at0) const
    {
        if (small) return ak(a.ren0_x, a.ren0_y, a.nren0, at0);
        return tab(S40, sqrt(r2)) - bicubic(x, y);
    }
};

// geometry of the pair (functions.c:1498-1503, 3472-3479) and the values at the two
// end points (z(chi_i) and the biases there)
struct pt {
    double mu, chi1, chi2, c, omc, sint;
    double s1, s2, b1, b2, fe1, fe2;
    double Dc1, Dc2, fc1, fc2, cHc1, cHc2, ac1, ac2, G1c, G2c, G1c2;
};

__device__ pt make_pt(const intg_args &a, double zm, double mu, double sep)
{
    const lk L(a);
    pt p;
    p.mu = mu;
    const double cm = ak(a.bg_z, a.bg_chi, a.nbg, zm);
    p.chi1 = cm - sep * mu / 2;
    p.chi2 = cm + sep * mu / 2;
    const double den = 2 * cm * cm - mu * mu * sep * sep / 2;
    p.c = (2 * cm * cm - sep * sep + mu * mu * sep * sep / 2) / den;
    // 1 - cos(theta) without cancellation
    p.omc = sep * sep * (1 - mu * mu) / den;
    p.sint = sqrt(fmax(p.omc * (1 + p.c), 0.0));
    const double z1c = L.z_of(p.chi1), z2c = L.z_of(p.chi2);
    p.s1 = ak(a.mb_z, a.s1, a.nmb, z1c);
    p.s2 = ak(a.mb_z, a.s2, a.nmb, z2c);
    p.fe1 = ak(a.mb_z, a.fe1, a.nmb, z1c);
    p.fe2 = ak(a.mb_z, a.fe2, a.nmb, z2c);
    p.b1 = ak(a.bias_z, a.b1, a.nb, z1c);
    p.b2 = ak(a.bias_z, a.b2, a.nb, z2c);
    p.Dc1 = L.bgc(C_D1, z1c);
    p.Dc2 = L.bgc(C_D1, z2c);
    p.fc1 = L.bgc(C_F, z1c);
    p.fc2 = L.bgc(C_F, z2c);
    p.cHc1 = L.bgc(C_CH, z1c);
    p.cHc2 = L.bgc(C_CH, z2c);
    p.ac1 = L.bgc(C_A, z1c);
    p.ac2 = L.bgc(C_A, z2c);
    p.G1c = L.bg(a.bg_G1, z1c);
    p.G2c = L.bg(a.bg_G2, z2c);
    p.G1c2 = L.bg(a.bg_G1, z2c);
    return p;
}

// The values along line of sight 1 that the double-integrated terms read at one x1:
// D1, a, conformal H and f at z(chi1 x1) (the last two with g5 only)
enum { X1_D1, X1_A, X1_CH, X1_F, X1_N };

__device__ void x1_values(const intg_args &a, double l1, double *v)
{
    const lk L(a);
    const double z1 = L.z_of(l1);
    const int j1 = L.bg_interval(z1);
    v[X1_D1] = L.bgcj(C_D1, j1, z1);
    v[X1_A] = L.bgcj(C_A, j1, z1);
    v[X1_CH] = v[X1_F] = 0.0;
    if (a.flags & F_G5) {
        v[X1_CH] = L.bgcj(C_CH, j1, z1);
        v[X1_F] = L.bgcj(C_F, j1, z1);
    }
}

// functions_double_integrated (functions.c:3460-4017) at (x1, x2), times (H0 h)^4; v1
// holds x1_values at chi1 x1
__device__ __forceinline__ double double_point(const intg_args &a, const pt &p, double x1, double x2,
                                               const double *v1)
{
    const lk L(a);
    const unsigned f = a.flags;
    const double c = p.c, Om2 = a.Om * a.Om;
    const double l1 = p.chi1 * x1, l2 = p.chi2 * x2;
    const double r2 = fmax(l1 * l1 + l2 * l2 - 2 * l1 * l2 * c, 0.0);
    const double r = sqrt(r2);
    const double z2 = L.z_of(l2);
    int j2;
    const double D1 = v1[X1_D1], D2 = L.bgc(C_D1, z2, &j2);
    const double DD = D1 * D2 / v1[X1_A] / L.bgcj(C_A, j2, z2);
    const double LL = l1 * l2;
    const double q2 = (2 - 5 * p.s1) * (2 - 5 * p.s2);
    const bool lens = f & F_LEN, g4 = f & F_G4, g5 = f & F_G5, oc = f & F_ONLY_CROSS;
    double xr = 0.0;
    const int jr = L.tab_interval(r, &xr);
    double res = 0.0;
    double ren = 0.0;
    if (g4 || g5) ren = L.ren(r2, l1, l2, r2 <= 1e-12, l1);
    if (lens && !oc && !(f & F_FLAT_LL)) {
        double T;
        if (r2 > 1e-20) {
            const double cc1 = c * c - 1;
            T = 2 * cc1 * LL * L.tabj(S00, jr, xr, r) / 5 + 4 * c * L.tabj(S20, jr, xr, r) / 3 +
                4 * c * (r2 + 6 * c * LL) * L.tabj(S11, jr, xr, r) / 15 +
                2 * cc1 * LL * (2 * r2 + 3 * c * LL) * L.tabj(S02, jr, xr, r) / 7 / r2 +
                2 * c * (2 * r2 * r2 + 12 * c * r2 * LL + 15 * cc1 * LL * LL) * L.tabj(S13, jr, xr, r) / 15 / r2 +
                cc1 * LL * (6 * r2 * r2 + 30 * c * r2 * LL + 35 * cc1 * LL * LL) * L.tabj(S04, jr, xr, r) / 35 / r2 / r2;
        } else {
            T = 4 + 24 * LL * L.tab(S11, 0.0) / 15;
        }
        res += 9 * Om2 * q2 / 4 * p.chi1 * p.chi2 * DD * (1 - x1) * (1 - x2) * T;
    }
    double cH1 = 0, cH2 = 0, f1 = 0, f2 = 0;
    if (g5) {
        cH1 = v1[X1_CH];
        cH2 = L.bgcj(C_CH, j2, z2);
        f1 = v1[X1_F];
        f2 = L.bgcj(C_F, j2, z2);
    }
    if (g4 && !oc) res += 9 * Om2 * q2 * DD * ren;
    if (g5 && !oc) res += 9 * Om2 * p.G1c * p.G2c * p.chi1 * p.chi2 * DD * cH1 * cH2 * (f1 - 1) * (f2 - 1) * ren;
    if ((g4 || g5) && lens) {
        const double P = r2 != 0.0 ? 2 * LL * c * L.tabj(S31, jr, xr, r) - LL * LL * (1 - c * c) * L.tabj(S22, jr, xr, r)
                                   : 2 * LL * L.tab(S31, 0.0);
        const double w2 = (1 - x2) / x2, w1 = (1 - x1) / x1;
        if (g4) res += 9 * Om2 / 2 * q2 * DD * (w2 + w1) * P;
        // functions.c:3927-3940: conformal_Hz (conformal_Hz - 1) as written there
        if (g5)
            res += 9 * Om2 / 2 * DD * P *
                   ((2 - 5 * p.s2) * p.G1c * p.chi1 * cH1 * (cH1 - 1) * w2 +
                    (2 - 5 * p.s1) * p.G2c * p.chi2 * cH2 * (cH2 - 1) * w1);
    }
    if (g4 && g5)
        res += 9 * Om2 * DD * ren *
               (p.G2c * (2 - 5 * p.s1) * p.chi2 * cH2 * (f2 - 1) + p.G1c * (2 - 5 * p.s2) * p.chi1 * cH1 * (f1 - 1));
    const double h2 = a.H0h * a.H0h;
    return res * h2 * h2;
}

// The numerators of the single-integrated lensing terms that lose digits near the
// line of closest approach (functions.c:2183-2187, 2296-2303, 2638-2644; chi the local
// end point's distance, lambda the integration point's), in d = chi - lambda and
// e = 1 - cos(theta), where the squared separation is d^2 + 2 chi lambda e; each is the
// reference's polynomial identically, without its FP64 cancellation.
__device__ inline double num_rsd_i02(double ch, double lam, double d, double e)
{
    return 2 * d * d * (4 * d - ch) + 2 * e * (ch * ch * ch - 11 * ch * ch * d + 19 * ch * d * d - 12 * d * d * d) -
           3 * e * e * lam * (ch * ch - 2 * ch * d + 4 * d * d) - 6 * ch * e * e * e * lam * lam;
}

__device__ inline double num_rsd_i04(double ch, double lam, double d, double e)
{
    const double d2 = d * d, ch2 = ch * ch, e2 = e * e, lm2 = lam * lam;
    return 4 * d2 * d2 * (d - 2 * ch) + 4 * d2 * e * (12 * ch2 * ch - 23 * ch2 * d + 15 * ch * d2 - 3 * d2 * d) +
           6 * e2 * lm2 * (-2 * ch2 * ch + 6 * ch2 * d - 6 * ch * d2 + d2 * d) +
           2 * ch * e2 * e * lm2 * (2 * ch2 - 4 * ch * d + 3 * d2) + ch2 * e2 * e2 * lm2 * lam;
}

__device__ inline double num_d1_i13(double ch, double lam, double d, double e)
{
    const double d2 = d * d, ch2 = ch * ch;
    return 2 * d2 * d * (3 * ch - d) + 2 * d * e * (-9 * ch2 * ch + 14 * ch2 * d - 8 * ch * d2 + d2 * d) -
           ch * e * e * lam * (2 * ch2 - 5 * ch * d + 4 * d2) - ch2 * e * e * e * lam * lam;
}

// functions_single_integrated (functions.c:1486-3458) at x: point 1 at chi1 with the
// integrated source along line of sight 2 (lambda2 = x chi2, squared separation r21)
// and the other way round (r22); the branches where neither is exactly zero
__device__ double single_point(const intg_args &a, const pt &p, double x)
{
    const lk L(a);
    const unsigned f = a.flags;
    const double c = p.c, Om = a.Om, H2 = a.H0h * a.H0h, H4 = H2 * H2;
    const double chi1 = p.chi1, chi2 = p.chi2;
    const double l1 = chi1 * x, l2 = chi2 * x;
    const double r21 = fmax(l2 * l2 + chi1 * chi1 - 2 * chi1 * l2 * c, 0.0);
    const double r22 = fmax(l1 * l1 + chi2 * chi2 - 2 * chi2 * l1 * c, 0.0);
    const double rA = sqrt(r21), rB = sqrt(r22);
    double xA = 0.0, xB = 0.0;
    const int jA = L.tab_interval(rA, &xA), jB = L.tab_interval(rB, &xB);
    const double z1 = L.z_of(l1), z2 = L.z_of(l2);
    const int j1 = L.bg_interval(z1), j2 = L.bg_interval(z2);
    const double D1x = L.bgcj(C_D1, j1, z1), D2x = L.bgcj(C_D1, j2, z2);
    const double a1x = L.bgcj(C_A, j1, z1), a2x = L.bgcj(C_A, j2, z2);
    const double VA = D2x / a2x, VB = D1x / a1x;
    const double WA = (1 - x) * VA, WB = (1 - x) * VB;
    const double s1 = p.s1, s2 = p.s2, c2 = 2 * c * c - 1, e = p.omc;
    const bool lens = f & F_LEN, g4 = f & F_G4, g5 = f & F_G5, full_dl = !(f & F_FLAT_DL);
    double res = 0.0;
    if (lens && (f & F_DEN) && full_dl)   // functions.c:1707-1781
        res += H2 * (-3 * Om / 2) *
               (p.b1 * (2 - 5 * s2) * p.Dc1 * chi2 * WA *
                    (2 * chi1 * c * L.tabj(S11, jA, xA, rA) - chi1 * chi1 * l2 * (1 - c * c) * L.tabj(S02, jA, xA, rA) / r21) +
                p.b2 * (2 - 5 * s1) * p.Dc2 * chi1 * WB *
                    (2 * chi2 * c * L.tabj(S11, jB, xB, rB) - chi2 * chi2 * l1 * (1 - c * c) * L.tabj(S02, jB, xB, rB) / r22));
    if (lens && (f & F_RSD) && full_dl) {   // functions.c:2167-2198, then 2291-2324 or 2377-2390
        const bool wide = fabs(p.mu) < 0.999;
        auto rsd = [&](double ch, double lam, double rr, double rv, int jv, double xv) {
            (void)rr;
            const double d = ch - lam, rs = d * d + 2 * ch * lam * e;
            double v = (lam - 6 * ch * c + 3 * lam * c2) * L.tabj(S00, jv, xv, rv) / 15 -
                       num_rsd_i02(ch, lam, d, e) * L.tabj(S02, jv, xv, rv) / rs / 21;
            if (wide) v -= num_rsd_i04(ch, lam, d, e) * L.tabj(S04, jv, xv, rv) / rs / rs / 35;
            else v += 4 * (lam + ch) * L.tabj(S04, jv, xv, rv) / 35;
            return v;
        };
        res += H2 * 3 * Om / 2 *
               (chi2 * p.fc1 * (2 - 5 * s2) * p.Dc1 * WA * rsd(chi1, l2, r21, rA, jA, xA) +
                chi1 * p.fc2 * (2 - 5 * s1) * p.Dc2 * WB * rsd(chi2, l1, r22, rB, jB, xB));
    }
    if (lens && (f & F_D1) && full_dl) {   // functions.c:2624-2661 (G1 at z(chi2), 2 - 5 s2 in the second half)
        auto d1k = [&](double ch, double lam, double rr, double rv, int jv, double xv) {
            (void)rr;
            const double d = ch - lam, rs = d * d + 2 * ch * lam * e;
            return 2 * (c * (lam * lam - 2 * ch * ch) + ch * lam * (2 * c2 - 1)) * L.tabj(S11, jv, xv, rv) / 15 +
                   2 * c * L.tabj(S20, jv, xv, rv) / 3 - num_d1_i13(ch, lam, d, e) * L.tabj(S13, jv, xv, rv) / rs / 15;
        };
        res += H2 * 3 * Om / 2 *
               (chi2 * p.cHc1 * p.fc1 * p.G1c * (2 - 5 * s2) * p.Dc1 * WA * d1k(chi1, l2, r21, rA, jA, xA) +
                chi1 * p.cHc2 * p.fc2 * p.G1c2 * (2 - 5 * s2) * p.Dc2 * WB * d1k(chi2, l1, r22, rB, jB, xB));
    }
    if (lens && (f & (F_D2 | F_G1 | F_G2 | F_G3))) {
        const double PA = 2 * chi1 * c * L.tabj(S31, jA, xA, rA) - chi1 * chi1 * l2 * (1 - c * c) * L.tabj(S22, jA, xA, rA);
        const double PB = 2 * chi2 * c * L.tabj(S31, jB, xB, rB) - chi2 * chi2 * l1 * (1 - c * c) * L.tabj(S22, jB, xB, rB);
        const double A = chi2 * p.Dc1 * WA * PA, B = chi1 * p.Dc2 * WB * PB;
        const double q = 9 * Om * Om / 4 * H4;
        if (f & F_D2)   // functions.c:3000-3030
            res += H2 * (-3 * Om / 2) *
                   ((3 - p.fe1) * p.fc1 * p.cHc1 * p.cHc1 * (2 - 5 * s2) * A +
                    (3 - p.fe2) * p.fc2 * p.cHc2 * p.cHc2 * (2 - 5 * s1) * B);
        if (f & F_G1) res += q * ((1 + p.G1c) * (2 - 5 * s2) * A + (1 + p.G2c) * (2 - 5 * s1) * B);   // 3032-3060
        if (f & F_G2) res += q * ((5 * s1 - 2) * (2 - 5 * s2) * A + (5 * s2 - 2) * (2 - 5 * s1) * B);  // 3062-3090
        if (f & F_G3) res += q * ((p.fc1 - 1) * (2 - 5 * s2) * A + (p.fc2 - 1) * (2 - 5 * s1) * B);   // 3092-3120
    }
    if (g4 || g5) {
        const bool div = f & (F_D2 | F_G1 | F_G2 | F_G3 | F_G4 | F_G5);
        const double ren1 = div ? L.ren(r21, l2, chi1, r21 == 0.0, l2) : 0.0;
        const double ren2 = div ? L.ren(r22, l1, chi2, r22 == 0.0, l1) : 0.0;
        const double E1 = L.bgcj(C_CH, j1, z1) * (L.bgcj(C_F, j1, z1) - 1),
                     E2 = L.bgcj(C_CH, j2, z2) * (L.bgcj(C_F, j2, z2) - 1);
        auto rsdk = [&](double rr, double lam, double rv, int jv, double xv) {
            return (2 * rr / 3 + (c * c - 1) * lam * lam) * L.tabj(S22, jv, xv, rv) - L.tabj(S20, jv, xv, rv) / 3;
        };
        const double q = 9 * Om * Om / 2 * H4;
        const double cH2c1 = p.cHc1 * p.cHc1, cH2c2 = p.cHc2 * p.cHc2;
        if (g4) {
            if (f & F_DEN)   // functions.c:3122-3140
                res += H2 * (-3 * Om) *
                       (p.b1 * (2 - 5 * s2) * p.Dc1 * VA * L.tabj(S20, jA, xA, rA) + p.b2 * (2 - 5 * s1) * p.Dc2 * VB * L.tabj(S20, jB, xB, rB));
            if (f & F_RSD)   // functions.c:3164-3189
                res += H2 * 3 * Om *
                       (p.fc1 * (2 - 5 * s2) * p.Dc1 * VA * rsdk(r21, l2, rA, jA, xA) +
                        p.fc2 * (2 - 5 * s1) * p.Dc2 * VB * rsdk(r22, l1, rB, jB, xB));
            if (f & F_D1)    // functions.c:3220-3235
                res += H2 * 3 * Om *
                       (p.cHc1 * p.fc1 * (2 - 5 * s2) * p.Dc1 * VA * (l2 * c - chi1) * L.tabj(S31, jA, xA, rA) +
                        p.cHc2 * p.fc2 * (2 - 5 * s1) * p.Dc2 * VB * (l1 * c - chi2) * L.tabj(S31, jB, xB, rB));
            if (f & F_D2)    // functions.c:3258-3275
                res += H2 * (-3 * Om) *
                       ((3 - p.fe1) * p.fc1 * cH2c1 * (2 - 5 * s2) * p.Dc1 * VA * ren1 +
                        (3 - p.fe2) * p.fc2 * cH2c2 * (2 - 5 * s1) * p.Dc2 * VB * ren2);
            const double u1 = p.Dc1 / p.ac1 * VA * ren1, u2 = p.Dc2 / p.ac2 * VB * ren2;
            if (f & F_G1) res += q * ((1 + p.G1c) * (2 - 5 * s2) * u1 + (1 + p.G2c) * (2 - 5 * s1) * u2);   // 3298-3315
            if (f & F_G2) res += q * ((5 * s1 - 2) * (2 - 5 * s2) * u1 + (5 * s2 - 2) * (2 - 5 * s1) * u2);  // 3340-3357
            if (f & F_G3) res += q * ((p.fc1 - 1) * (2 - 5 * s2) * u1 + (p.fc2 - 1) * (2 - 5 * s1) * u2);   // 3382-3399
        }
        if (g5) {
            if (f & F_DEN)   // functions.c:3142-3162 (D1 a along the line of sight, as written there)
                res += H2 * (-3 * Om) *
                       (chi2 * p.b1 * p.G2c * p.Dc1 * E2 * D2x * a2x * L.tabj(S20, jA, xA, rA) +
                        chi1 * p.b2 * p.G1c * p.Dc2 * E1 * D1x * a1x * L.tabj(S20, jB, xB, rB));
            if (f & F_RSD)   // functions.c:3191-3218
                res += H2 * 3 * Om *
                       (chi2 * p.fc1 * p.G2c * p.Dc1 * E2 * VA * rsdk(r21, l2, rA, jA, xA) +
                        chi1 * p.fc2 * p.G1c * p.Dc2 * E1 * VB * rsdk(r22, l1, rB, jB, xB));
            if (f & F_D1)    // functions.c:3237-3256
                res += H2 * 3 * Om *
                       (chi2 * p.cHc1 * p.fc1 * p.G2c * p.Dc1 * E2 * VA * (l2 * c - chi1) * L.tabj(S31, jA, xA, rA) +
                        chi1 * p.cHc2 * p.fc2 * p.G1c * p.Dc2 * E1 * VB * (l1 * c - chi2) * L.tabj(S31, jB, xB, rB));
            if (f & F_D2)    // functions.c:3277-3296
                res += H2 * (-3 * Om) *
                       (chi2 * (3 - p.fe1) * p.fc1 * cH2c1 * p.G2c * p.Dc1 * E2 * VA * ren1 +
                        chi1 * (3 - p.fe2) * p.fc2 * cH2c2 * p.G1c * p.Dc2 * E1 * VB * ren2);
            const double u1 = chi2 * p.Dc1 / p.ac1 * E2 * VA * ren1, u2 = chi1 * p.Dc2 / p.ac2 * E1 * VB * ren2;
            if (f & F_G1) res += q * ((1 + p.G1c) * p.G2c * u1 + (1 + p.G2c) * p.G1c * u2);   // 3317-3338
            if (f & F_G2) res += q * ((5 * s1 - 2) * p.G2c * u1 + (5 * s2 - 2) * p.G1c * u2);  // 3359-3380
            if (f & F_G3) res += q * ((p.fc1 - 1) * p.G2c * u1 + (p.fc2 - 1) * p.G1c * u2);   // 3401-3422
        }
    }
    return res;
}

// node q of the composite rule on [lo, hi]: `panels` equal panels of the n-point
// Gauss-Legendre rule (gx, gw on [-1, 1]); returns the node, weight in *w
__device__ inline double composite(double lo, double hi, int panels, int n, const double *gx, const double *gw, int q,
                                   double *w)
{
    const int k = q / n, j = q - k * n;
    const double step = (hi - lo) / panels;
    const double e0 = lo + k * step, e1 = (k + 1 == panels) ? hi : lo + (k + 1) * step;
    const double h = (e1 - e0) / 2;
    *w = h * gw[j];
    return (e0 + e1) / 2 + h * gx[j];
}

// the pieces of the single-integrated rule: [a, b] mapped about centre cc with width s
struct piece {
    double a, b, cc, s;
};

__device__ int single_pieces(const pt &p, piece *pc)
{
    const double cA = p.chi1 * p.c / p.chi2, sA = fmax(p.chi1 * p.sint / p.chi2, 1e-12);
    const double cB = p.chi2 * p.c / p.chi1, sB = fmax(p.chi2 * p.sint / p.chi1, 1e-12);
    double in[2];
    int nin = 0;
    if (cA > 0.0 && cA < 1.0) in[nin++] = cA;
    if (cB > 0.0 && cB < 1.0 && !(nin == 1 && cB == cA)) in[nin++] = cB;
    if (nin == 2 && in[1] < in[0]) {
        const double t = in[0];
        in[0] = in[1];
        in[1] = t;
    }
    double cuts[5];
    int nc = 0;
    cuts[nc++] = 0.0;
    if (nin == 2) {
        cuts[nc++] = in[0];
        cuts[nc++] = (in[0] + in[1]) / 2;
        cuts[nc++] = in[1];
    } else if (nin == 1) {
        cuts[nc++] = in[0];
    }
    cuts[nc++] = 1.0;
    int np = 0;
    for (int i = 0; i + 1 < nc; ++i) {
        const double a = cuts[i], b = cuts[i + 1];
        if (b <= a) continue;
        const double m = (a + b) / 2;
        const double dA = fmin(fmin(fabs(cA - m), fabs(cA - a)), fabs(cA - b));
        const double dB = fmin(fmin(fabs(cB - m), fabs(cB - a)), fabs(cB - b));
        double cc = dA <= dB ? cA : cB;
        const double s = dA <= dB ? sA : sB;
        if (a < cc && cc < b) cc = fabs(cc - a) < fabs(cc - b) ? a : b;
        pc[np++] = piece{a, b, cc, s};
    }
    return np;
}

// the value records and node pairs of up to 16 curves, one thread per (node, curve)
struct coef_list {
    int ncurve;
    const double *x[16], *y[16];
    int n[16];
    double *val[16], *pair[16];
};

__global__ void __launch_bounds__(256) k_intg_coef(coef_list L)
{
    const int t = blockIdx.y;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = L.n[t];
    if (j >= n) return;
    const double *x = L.x[t], *y = L.y[t];
    double v[VAL] = {y[j], 0.0, 0.0, 0.0}, x1 = x[j];   // the last node's: no segment
    if (j < n - 1) {
        const ce::akima_cf<double> cf = ce::akima_coefs<double>(x, nullptr, y, nullptr, n, j);
        v[1] = cf.b;
        v[2] = cf.c;
        v[3] = cf.d;
        x1 = x[j + 1];
    }
    double *r = L.val[t] + (size_t)VAL * j;
    *reinterpret_cast<double2 *>(r) = double2{v[0], v[1]};
    *reinterpret_cast<double2 *>(r + 2) = double2{v[2], v[3]};
    *reinterpret_cast<double2 *>(L.pair[t] + (size_t)PAIR * j) = double2{x[j], x1};
}

// one workgroup per (z, mu, r): PART 0 sums the single-integrated terms into out[..][0],
// PART 1 the double-integrated ones into out[..][1]; two kernels, so that each gets the
// registers of its own part (together they held 252 VGPRs, two wavefronts per SIMD).
// The arguments live in device memory (a pointer, not a by-value struct, whose address
// the lookups would otherwise take in private memory).
template <int PART>
__device__ __forceinline__ void integrated_body(const intg_args *__restrict__ pa)
{
    const intg_args &a = *pa;
    const int idx = blockIdx.x;   // (z, mu, r), r fastest
    const int ir = idx % a.nsep;
    const int im = (idx / a.nsep) % a.nmu;
    const int iz = idx / (a.nsep * a.nmu);
    const pt p = make_pt(a, a.zmean[iz], a.mu[im], a.sep[ir]);
    const unsigned f = a.flags;
    double acc_s = 0.0, acc_d = 0.0;
    const bool want_single =
        ((f & F_LEN) && (f & (F_DEN | F_RSD | F_D1 | F_D2 | F_G1 | F_G2 | F_G3))) ||
        ((f & (F_G4 | F_G5)) && (f & (F_DEN | F_RSD | F_D1 | F_D2 | F_G1 | F_G2 | F_G3)));
    if (PART == 0 && want_single) {
        // the pieces in LDS (a private array indexed by node would live in scratch)
        __shared__ piece pc[4];
        __shared__ int npc;
        __shared__ pt sp;
        if (threadIdx.x == 0) {
            sp = p;
            piece t[4];
            npc = single_pieces(p, t);
            for (int i = 0; i < 4; ++i) pc[i] = t[i];
        }
        __syncthreads();
        const int np = npc;
        const int per = a.ps * a.ns;
        for (int q = threadIdx.x; q < np * per; q += blockDim.x) {
            const piece &P = pc[q / per];
            const double sgn = P.a >= P.cc ? 1.0 : -1.0;
            const double ua = asinh(sgn * (P.a - P.cc) / P.s), ub = asinh(sgn * (P.b - P.cc) / P.s);
            double w;
            const double u = composite(fmin(ua, ub), fmax(ua, ub), a.ps, a.ns, a.gsx, a.gsw, q % per, &w);
            const double x = P.cc + sgn * P.s * sinh(u);
            acc_s += w * P.s * cosh(u) * single_point(a, sp, x);
        }
    }
    if (PART == 1 && (f & (F_LEN | F_G4 | F_G5))) {
        const double xb = p.chi1 * p.c > p.chi2 ? p.chi2 / (p.chi1 * p.c) : 1.0;
        const int pA = a.p1, pB = xb < 1.0 ? (a.p1 / 2 > 1 ? a.p1 / 2 : 1) : 0;
        const int n1 = (pA + pB) * a.n1, nuu = a.pu * a.nu;
        // per x1 node (CE_INTG_MAX_X1 at most, checked by the host; the launch sizes the
        // dynamic LDS for the rule's count): the node, its weight, the sinh scale, the u
        // interval and x1_values, once per workgroup
        extern __shared__ double sxd[];
        double(*sx)[5 + X1_N] = reinterpret_cast<double(*)[5 + X1_N]>(sxd);
        for (int i1 = threadIdx.x; i1 < n1; i1 += blockDim.x) {
            double w1;
            const double x1 = i1 < pA * a.n1 ? composite(0.0, xb, pA, a.n1, a.g1x, a.g1w, i1, &w1)
                                             : composite(xb, 1.0, pB, a.n1, a.g1x, a.g1w, i1 - pA * a.n1, &w1);
            const double l1 = p.chi1 * x1, centre = l1 * p.c;
            const double sc = fmax(l1 * p.sint, 1e-9 * p.chi2);
            sx[i1][0] = x1;
            sx[i1][1] = w1;
            sx[i1][2] = sc;
            sx[i1][3] = asinh((0.0 - centre) / sc);
            sx[i1][4] = asinh((p.chi2 - centre) / sc);
            x1_values(a, l1, &sx[i1][5]);
        }
        __syncthreads();
        for (int q = threadIdx.x; q < n1 * nuu; q += blockDim.x) {
            const int i1 = q / nuu, iu = q - i1 * nuu;
            const double x1 = sx[i1][0], w1 = sx[i1][1], sc = sx[i1][2];
            double wu;
            const double u = composite(sx[i1][3], sx[i1][4], a.pu, a.nu, a.gux, a.guw, iu, &wu);
            // sinh and cosh from one exponential (lambda2's absolute error stays ~eps sc)
            const double eu = exp(u), ei = 1.0 / eu;
            const double lam2 = p.chi1 * x1 * p.c + sc * (0.5 * (eu - ei));
            acc_d += w1 * wu * sc * (0.5 * (eu + ei)) / p.chi2 * double_point(a, p, x1, lam2 / p.chi2, &sx[i1][5]);
        }
    }
    // workgroup sum: wavefront shuffles, then one slot per wavefront in LDS
    double acc = PART == 0 ? acc_s : acc_d;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
    __shared__ double red[4];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red[wv] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double ss = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) ss += red[k];
        a.out[2 * (int64_t)idx + PART] = ss;
    }
}

__global__ void __launch_bounds__(256) k_integrated_single(const intg_args *__restrict__ pa)
{
    integrated_body<0>(pa);
}

__global__ void __launch_bounds__(256) k_integrated_double(const intg_args *__restrict__ pa)
{
    integrated_body<1>(pa);
}

}  // namespace

extern "C" int ce_launch_integrated(void *stream, const ce_intg_dev *in, unsigned flags, int nz, const double *d_z,
                                    int nmu, const double *d_mu, int nsep, const double *d_sep, const int *rule,
                                    const double *d_gl, double *d_out)
{
    intg_args a;
    a.nbg = in->nbg;
    a.bg_z = in->bg_z;
    a.bg_chi = in->bg_chi;
    a.bg_D1 = in->bg_D1;
    a.bg_a = in->bg_a;
    a.bg_f = in->bg_f;
    a.bg_cH = in->bg_cH;
    a.bg_G1 = in->bg_G1;
    a.bg_G2 = in->bg_G2;
    a.nb = in->nb;
    a.bias_z = in->bias_z;
    a.b1 = in->b1;
    a.b2 = in->b2;
    a.nmb = in->nmb;
    a.mb_z = in->mb_z;
    a.s1 = in->s1;
    a.s2 = in->s2;
    a.fe1 = in->fe1;
    a.fe2 = in->fe2;
    a.tab_x = in->tab_x;
    a.tab_y = in->tab_y;
    for (int t = 0; t < NSLOT; ++t) {
        a.tab_off[t] = in->tab_off[t];
        a.tab_len[t] = in->tab_len[t];
        a.tab_f0[t] = in->tab_f0[t];
        a.tab_lx0[t] = in->tab_lx0[t];
        a.tab_lsc[t] = in->tab_lsc[t];
    }
    a.nguess = in->nguess;
    a.chi_guess = in->chi_guess;
    a.tab_ref = in->tab_ref;
    // the value records, then the node pairs; each the tables' in tab_y's layout, then
    // the background curves'
    const size_t nrec = (size_t)in->ntab_values + (size_t)NBGC * in->nbg;
    double *tval = in->coef, *bval = tval + (size_t)VAL * in->ntab_values;
    double *tpair = in->coef + (size_t)VAL * nrec, *bpair = tpair + (size_t)PAIR * in->ntab_values;
    a.tab_val = tval;
    a.bgc_val = bval;
    a.tab_pair = tpair;
    a.bgc_pair = bpair;
    coef_list cl;
    cl.ncurve = 0;
    int nmax = 0;
    for (int t = 0; t < NSLOT; ++t) {
        if (in->tab_len[t] < 5) continue;
        const int q = cl.ncurve++;
        cl.x[q] = in->tab_x + in->tab_off[t];
        cl.y[q] = in->tab_y + in->tab_off[t];
        cl.n[q] = in->tab_len[t];
        cl.val[q] = tval + (size_t)VAL * in->tab_off[t];
        cl.pair[q] = tpair + (size_t)PAIR * in->tab_off[t];
        nmax = cl.n[q] > nmax ? cl.n[q] : nmax;
    }
    const double *bx[NBGC] = {in->bg_z, in->bg_z, in->bg_z, in->bg_z, in->bg_chi};
    const double *by[NBGC] = {in->bg_D1, in->bg_a, in->bg_f, in->bg_cH, in->bg_z};
    for (int k = 0; k < NBGC; ++k) {
        const int q = cl.ncurve++;
        cl.x[q] = bx[k];
        cl.y[q] = by[k];
        cl.n[q] = in->nbg;
        cl.val[q] = bval + (size_t)VAL * k * in->nbg;
        cl.pair[q] = bpair + (size_t)PAIR * k * in->nbg;
        nmax = in->nbg > nmax ? in->nbg : nmax;
    }
    hipLaunchKernelGGL(k_intg_coef, dim3((unsigned)((nmax + 255) / 256), (unsigned)cl.ncurve), dim3(256), 0,
                       (hipStream_t)stream, cl);
    a.nren0 = in->nren0;
    a.ren0_x = in->ren0_x;
    a.ren0_y = in->ren0_y;
    a.nren = in->nren;
    a.ren_chi = in->ren_chi;
    a.ren_rec = in->ren_rec;
    a.Om = in->Om;
    a.H0h = in->H0h;
    a.flags = flags;
    a.nz = nz;
    a.nmu = nmu;
    a.nsep = nsep;
    a.zmean = d_z;
    a.mu = d_mu;
    a.sep = d_sep;
    a.p1 = rule[0];
    a.n1 = rule[1];
    a.pu = rule[2];
    a.nu = rule[3];
    a.ps = rule[4];
    a.ns = rule[5];
    a.g1x = d_gl;
    a.g1w = d_gl + a.n1;
    a.gux = d_gl + 2 * a.n1;
    a.guw = a.gux + a.nu;
    a.gsx = a.guw + a.nu;
    a.gsw = a.gsx + a.ns;
    a.out = d_out;
    const int64_t blocks = (int64_t)nz * nmu * nsep;
    if (blocks <= 0 || blocks > 0x7fffffff) return (int)hipErrorInvalidValue;
    static_assert(sizeof(intg_args) <= CE_INTG_ARGS_BYTES, "intg_args outgrew its device slot");
    static_assert(NBGC == CE_INTG_NBGC, "background curves with segment records");
    hipError_t e = hipMemcpyAsync(in->args_dev, &a, sizeof(a), hipMemcpyHostToDevice, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);   /* `a` is on this stack */
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(k_integrated_single, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       (const intg_args *)in->args_dev);
    const int p1b = a.p1 / 2 > 1 ? a.p1 / 2 : 1;
    const size_t x1_lds = (size_t)(a.p1 + p1b) * a.n1 * (5 + X1_N) * sizeof(double);
    hipLaunchKernelGGL(k_integrated_double, dim3((unsigned)blocks), dim3(256), x1_lds, (hipStream_t)stream,
                       (const intg_args *)in->args_dev);
    return (int)hipGetLastError();
}
