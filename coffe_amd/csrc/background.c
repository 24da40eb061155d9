/*
 * Background tables that feed the multipole weights (input provider of the hot
 * path; host C).  Follows coffe_background_init (background.c:301-697):
 *   grid        z_i = 15 i / (bins - 1)                          (background.c:437)
 *   H(z)        sqrt(Om (1+z)^3 + Og (1+z)^4 + Ode wint(z)) h H0  (background.c:457-461)
 *   conformal   a H, and its conformal-time derivative            (background.c:462-472)
 *   D1, f       growth ODE D'' = -3/2 (1 - w/(1+x)) D'/a + 3/2 x/(1+x) D/a^2 from
 *               a = 0.05 with D = 0.05, D' = 1; f = D' a / D; D normalised to 1 at
 *               z = 0                                            (background.c:214-233,
 *                                                                  474-486, 540-545)
 *   chi(z)      int_0^z dz / E / (h H0)                          (background.c:488-502)
 *   G1, G2      H'/H^2 + (2 - 5 s)/(chi H) + 5 s - f_evo, zero at z = 0
 *                                                                (background.c:504-536)
 * with w(z) = w0 + wa z/(1+z) (common.c:678-685).  The dark-energy and x(a)
 * integrals are taken in closed form for this w(z); chi by 61-point Gauss-Kronrod
 * per grid interval; the ODE by classical RK4 in ln a with 64 steps per grid
 * interval.  (The reference integrates with GSL's adaptive RK2 at 1e-6, so the
 * tables agree with it to about that level, not bit for bit.)
 *
 * Every quantity is carried as a value and its derivative along one direction in
 * the parameters (Omega0_cdm, Omega0_baryon, Omega0_gamma, w0, wa, h): the grid, the
 * RK4 steps and the quadrature nodes do not depend on the parameters, so the
 * derivative part is the exact derivative of the computed tables
 * (coffe_background_jvp); coffe_background_compute is the zero direction.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "coffe_background.h"

#define BG_H0 (1. / 2997.92458) /* h / Mpc */

/* a value and its derivative along the direction */
typedef struct { double v, d; } bd;

static bd bd_c(double v) { bd r = {v, 0.0}; return r; }
static bd bd_add(bd x, bd y) { bd r = {x.v + y.v, x.d + y.d}; return r; }
static bd bd_sub(bd x, bd y) { bd r = {x.v - y.v, x.d - y.d}; return r; }
static bd bd_mul(bd x, bd y) { bd r = {x.v * y.v, x.d * y.v + x.v * y.d}; return r; }
static bd bd_scale(double c, bd x) { bd r = {c * x.v, c * x.d}; return r; }
static bd bd_div(bd x, bd y)
{
    const double q = x.v / y.v;
    bd r = {q, (x.d - q * y.d) / y.v};
    return r;
}
static bd bd_exp(bd x)
{
    const double e = exp(x.v);
    bd r = {e, e * x.d};
    return r;
}
static bd bd_sqrt(bd x)
{
    const double s = sqrt(x.v);
    bd r = {s, 0.5 * x.d / s};
    return r;
}

typedef struct {
    bd Om, Og, Ode, w0, wa;
} bg_par;

static bd w_of_z(const bg_par *p, double z) { return bd_add(p->w0, bd_scale(z / (1 + z), p->wa)); }

/* exp(3 int_0^z (1 + w)/(1 + z') dz') */
static bd wint(const bg_par *p, double z)
{
    const double l = log1p(z);
    const bd e = bd_sub(bd_scale(l, bd_add(bd_c(1.0), bd_add(p->w0, p->wa))), bd_scale(z / (1 + z), p->wa));
    return bd_exp(bd_scale(3.0, e));
}

/* Om/(1 - Om) exp(-3 int_a^1 w(a')/a' da') with w(a) = w0 + wa (1 - a) */
static bd x_of_a(const bg_par *p, double a)
{
    const bd e = bd_add(bd_scale(log(a), bd_add(p->w0, p->wa)), bd_scale(1 - a, p->wa));
    return bd_mul(bd_div(p->Om, bd_sub(bd_c(1.0), p->Om)), bd_exp(bd_scale(3.0, e)));
}

static bd E_of_z(const bg_par *p, double z)
{
    const double zp = 1 + z;
    return bd_sqrt(bd_add(bd_add(bd_scale(zp * zp * zp, p->Om), bd_scale(zp * zp * zp * zp, p->Og)),
                          bd_mul(p->Ode, wint(p, z))));
}

/* y = (D, dD/da); derivative with respect to a */
static void growth_rhs(const bg_par *p, double a, const bd *y, bd *dy)
{
    const double z = 1. / a - 1;
    const bd w = w_of_z(p, z);
    const bd x = x_of_a(p, a);
    const bd opx = bd_add(bd_c(1.0), x);
    dy[0] = y[1];
    dy[1] = bd_add(bd_scale(-1.5 / a, bd_mul(bd_sub(bd_c(1.0), bd_div(w, opx)), y[1])),
                   bd_scale(1.5 / (a * a), bd_mul(bd_div(x, opx), y[0])));
}

/* RK4 in s = ln a from a0 to a1 */
static void growth_step(const bg_par *p, double a0, double a1, bd *y, int nsub)
{
    const double s0 = log(a0), s1 = log(a1);
    const double hs = (s1 - s0) / nsub;
    for (int i = 0; i < nsub; ++i) {
        const double s = s0 + i * hs;
        const double am = exp(s + 0.5 * hs), ae = exp(s + hs), a = exp(s);
        bd k1[2], k2[2], k3[2], k4[2], yt[2];
        growth_rhs(p, a, y, k1);
        for (int c = 0; c < 2; ++c) { k1[c] = bd_scale(a, k1[c]); yt[c] = bd_add(y[c], bd_scale(0.5 * hs, k1[c])); }
        growth_rhs(p, am, yt, k2);
        for (int c = 0; c < 2; ++c) { k2[c] = bd_scale(am, k2[c]); yt[c] = bd_add(y[c], bd_scale(0.5 * hs, k2[c])); }
        growth_rhs(p, am, yt, k3);
        for (int c = 0; c < 2; ++c) { k3[c] = bd_scale(am, k3[c]); yt[c] = bd_add(y[c], bd_scale(hs, k3[c])); }
        growth_rhs(p, ae, yt, k4);
        for (int c = 0; c < 2; ++c) {
            k4[c] = bd_scale(ae, k4[c]);
            const bd sum = bd_add(bd_add(k1[c], bd_scale(2.0, k2[c])), bd_add(bd_scale(2.0, k3[c]), k4[c]));
            y[c] = bd_add(y[c], bd_scale(hs / 6, sum));
        }
    }
}

/* 61-point Gauss-Kronrod (Kronrod part only) of 1/E over [z0, z1] */
#include "gk61_nodes.h"
static bd inv_E_integral(const bg_par *p, double z0, double z1)
{
    const double c = 0.5 * (z0 + z1), h = 0.5 * (z1 - z0);
    bd s = bd_scale(GK61K_W[30], bd_div(bd_c(1.0), E_of_z(p, c)));
    for (int j = 0; j < 30; ++j) {
        const double d = h * GK61K_X[j];
        s = bd_add(s, bd_scale(GK61K_W[j], bd_add(bd_div(bd_c(1.0), E_of_z(p, c - d)),
                                                   bd_div(bd_c(1.0), E_of_z(p, c + d)))));
    }
    return bd_scale(h, s);
}

static double lin_interp(const double *x, const double *y, int n, double v)
{
    if (n <= 0) return 0.0;
    if (v <= x[0]) return y[0];
    if (v >= x[n - 1]) return y[n - 1];
    int lo = 0, hi = n - 1;
    while (hi > lo + 1) {
        const int m = (lo + hi) / 2;
        if (x[m] > v) hi = m;
        else lo = m;
    }
    const double t = (v - x[lo]) / (x[lo + 1] - x[lo]);
    return y[lo] + t * (y[lo + 1] - y[lo]);
}

#define BG_NT 10   /* z, a, Hz, conformal_Hz, conformal_Hz_prime, D1, f, chi, G1, G2 */

/* out[t] and tout[t] (may be NULL) of length bins; dir = derivative direction in
   (Omega0_cdm, Omega0_baryon, Omega0_gamma, w0, wa, h) */
static int background_dual(const double par[6], const double dir[6], int bins, int n_bias, const double *bias_z,
                           const double *s1, const double *s2, const double *fevo1, const double *fevo2,
                           double *const *out, double *const *tout)
{
    if (bins < 2 || !(par[5] > 0)) return 1;
    bd *t[BG_NT];
    bd *mem = (bd *)malloc(sizeof(bd) * BG_NT * (size_t)bins);
    if (!mem) return 2;
    for (int q = 0; q < BG_NT; ++q) t[q] = mem + (size_t)q * bins;
    bd *z = t[0], *a = t[1], *Hz = t[2], *cH = t[3], *cHp = t[4], *D1 = t[5], *f = t[6], *chi = t[7], *G1 = t[8],
       *G2 = t[9];
    bg_par p;
    const bd Oc = {par[0], dir[0]}, Ob = {par[1], dir[1]};
    p.Om = bd_add(Oc, Ob);
    p.Og.v = par[2]; p.Og.d = dir[2];
    p.Ode = bd_sub(bd_c(1.0), bd_add(p.Om, p.Og));
    p.w0.v = par[3]; p.w0.d = dir[3];
    p.wa.v = par[4]; p.wa.d = dir[4];
    const bd hH0 = {par[5] * BG_H0, dir[5] * BG_H0};
    for (int i = 0; i < bins; ++i) {
        const double zi = 15. * i / (double)(bins - 1);
        const double zp = 1 + zi;
        const bd w = w_of_z(&p, zi);
        const bd wi = wint(&p, zi);
        z[i] = bd_c(zi);
        a[i] = bd_c(1. / zp);
        Hz[i] = bd_mul(E_of_z(&p, zi), hH0);
        cH[i] = bd_scale(a[i].v, Hz[i]);
        const bd num = bd_add(bd_scale(zp * zp * zp, bd_add(bd_scale(2 * zp, p.Og), p.Om)),
                              bd_mul(bd_add(bd_c(1.0), bd_scale(3.0, w)), bd_mul(p.Ode, wi)));
        cHp[i] = bd_mul(bd_scale(-0.5 / (zp * zp), num), bd_mul(hH0, hH0));
    }
    /* comoving distance, cumulative */
    chi[0] = bd_c(0.0);
    for (int i = 1; i < bins; ++i) chi[i] = bd_add(chi[i - 1], bd_div(inv_E_integral(&p, z[i - 1].v, z[i].v), hH0));
    /* growth: integrate from a = 0.05 up through the grid (descending z) */
    bd y[2] = {{0.05, 0.0}, {1.0, 0.0}};
    double acur = 0.05;
    for (int i = bins - 1; i >= 0; --i) {
        if (a[i].v > acur) {
            const int nsub = 64 + (int)(64 * fabs(log(a[i].v / acur)) / 0.01);
            growth_step(&p, acur, a[i].v, y, nsub > 4096 ? 4096 : nsub);
            acur = a[i].v;
        }
        D1[i] = y[0];
        f[i] = bd_div(bd_scale(a[i].v, y[1]), y[0]);
    }
    const bd D10 = D1[0];
    for (int i = 0; i < bins; ++i) D1[i] = bd_div(D1[i], D10);
    for (int i = 0; i < bins; ++i) {
        if (z[i].v > 1e-10) {
            const double sv1 = lin_interp(bias_z, s1, n_bias, z[i].v);
            const double sv2 = lin_interp(bias_z, s2, n_bias, z[i].v);
            const double e1 = lin_interp(bias_z, fevo1, n_bias, z[i].v);
            const double e2 = lin_interp(bias_z, fevo2, n_bias, z[i].v);
            const bd base = bd_div(cHp[i], bd_mul(cH[i], cH[i]));
            const bd ichiH = bd_div(bd_c(1.0), bd_mul(chi[i], cH[i]));
            G1[i] = bd_add(base, bd_add(bd_scale(2 - 5 * sv1, ichiH), bd_c(5 * sv1 - e1)));
            G2[i] = bd_add(base, bd_add(bd_scale(2 - 5 * sv2, ichiH), bd_c(5 * sv2 - e2)));
        } else {
            G1[i] = bd_c(0.0);
            G2[i] = bd_c(0.0);
        }
    }
    for (int q = 0; q < BG_NT; ++q)
        for (int i = 0; i < bins; ++i) {
            if (out && out[q]) out[q][i] = t[q][i].v;
            if (tout && tout[q]) tout[q][i] = t[q][i].d;
        }
    free(mem);
    return 0;
}

int coffe_background_compute(double Omega0_cdm, double Omega0_baryon, double Omega0_gamma, double w0, double wa,
                             double h, int bins, int n_bias, const double *bias_z, const double *s1, const double *s2,
                             const double *fevo1, const double *fevo2, double *z, double *a, double *Hz,
                             double *conformal_Hz, double *conformal_Hz_prime, double *D1, double *f, double *chi,
                             double *G1, double *G2)
{
    const double par[6] = {Omega0_cdm, Omega0_baryon, Omega0_gamma, w0, wa, h};
    const double dir[6] = {0, 0, 0, 0, 0, 0};
    double *const out[BG_NT] = {z, a, Hz, conformal_Hz, conformal_Hz_prime, D1, f, chi, G1, G2};
    return background_dual(par, dir, bins, n_bias, bias_z, s1, s2, fevo1, fevo2, out, NULL);
}

int coffe_background_jvp(const double *par, const double *dir, int bins, int n_bias, const double *bias_z,
                         const double *s1, const double *s2, const double *fevo1, const double *fevo2, double *out,
                         double *tout)
{
    if (!par || !dir || !out || !tout) return 1;
    double *o[BG_NT], *to[BG_NT];
    for (int q = 0; q < BG_NT; ++q) {
        o[q] = out + (size_t)q * bins;
        to[q] = tout + (size_t)q * bins;
    }
    return background_dual(par, dir, bins, n_bias, bias_z, s1, s2, fevo1, fevo2, o, to);
}
