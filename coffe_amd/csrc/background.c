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
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "coffe_background.h"

#define BG_H0 (1. / 2997.92458) /* h / Mpc */

typedef struct {
    double Om, Og, Ode, w0, wa;
} bg_par;

static double w_of_z(const bg_par *p, double z) { return p->w0 + p->wa * z / (1 + z); }

/* exp(3 int_0^z (1 + w)/(1 + z') dz') */
static double wint(const bg_par *p, double z)
{
    const double l = log1p(z);
    return exp(3.0 * ((1 + p->w0 + p->wa) * l - p->wa * z / (1 + z)));
}

/* Om/(1 - Om) exp(-3 int_a^1 w(a')/a' da') with w(a) = w0 + wa (1 - a) */
static double x_of_a(const bg_par *p, double a)
{
    return p->Om / (1 - p->Om) * exp(3.0 * ((p->w0 + p->wa) * log(a) + p->wa * (1 - a)));
}

static double E_of_z(const bg_par *p, double z)
{
    const double zp = 1 + z;
    return sqrt(p->Om * zp * zp * zp + p->Og * zp * zp * zp * zp + p->Ode * wint(p, z));
}

/* y = (D, dD/da); derivative with respect to a */
static void growth_rhs(const bg_par *p, double a, const double *y, double *dy)
{
    const double z = 1. / a - 1;
    const double w = w_of_z(p, z);
    const double x = x_of_a(p, a);
    dy[0] = y[1];
    dy[1] = -1.5 * (1 - w / (1 + x)) * y[1] / a + 1.5 * x / (1 + x) * y[0] / (a * a);
}

/* RK4 in s = ln a from a0 to a1 */
static void growth_step(const bg_par *p, double a0, double a1, double *y, int nsub)
{
    const double s0 = log(a0), s1 = log(a1);
    const double hs = (s1 - s0) / nsub;
    for (int i = 0; i < nsub; ++i) {
        const double s = s0 + i * hs;
        double k1[2], k2[2], k3[2], k4[2], yt[2];
        double a = exp(s);
        growth_rhs(p, a, y, k1);
        k1[0] *= a; k1[1] *= a;
        double am = exp(s + 0.5 * hs);
        yt[0] = y[0] + 0.5 * hs * k1[0]; yt[1] = y[1] + 0.5 * hs * k1[1];
        growth_rhs(p, am, yt, k2);
        k2[0] *= am; k2[1] *= am;
        yt[0] = y[0] + 0.5 * hs * k2[0]; yt[1] = y[1] + 0.5 * hs * k2[1];
        growth_rhs(p, am, yt, k3);
        k3[0] *= am; k3[1] *= am;
        double ae = exp(s + hs);
        yt[0] = y[0] + hs * k3[0]; yt[1] = y[1] + hs * k3[1];
        growth_rhs(p, ae, yt, k4);
        k4[0] *= ae; k4[1] *= ae;
        y[0] += hs / 6 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0]);
        y[1] += hs / 6 * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1]);
    }
}

/* 61-point Gauss-Kronrod (Kronrod part only) of 1/E over [z0, z1] */
#include "gk61_nodes.h"
static double inv_E_integral(const bg_par *p, double z0, double z1)
{
    const double c = 0.5 * (z0 + z1), h = 0.5 * (z1 - z0);
    double s = GK61K_W[30] / E_of_z(p, c);
    for (int j = 0; j < 30; ++j) {
        const double d = h * GK61K_X[j];
        s += GK61K_W[j] * (1. / E_of_z(p, c - d) + 1. / E_of_z(p, c + d));
    }
    return s * h;
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

int coffe_background_compute(double Omega0_cdm, double Omega0_baryon, double Omega0_gamma, double w0, double wa,
                             double h, int bins, int n_bias, const double *bias_z, const double *s1, const double *s2,
                             const double *fevo1, const double *fevo2, double *z, double *a, double *Hz,
                             double *conformal_Hz, double *conformal_Hz_prime, double *D1, double *f, double *chi,
                             double *G1, double *G2)
{
    if (bins < 2 || !(h > 0)) return 1;
    bg_par p;
    p.Om = Omega0_cdm + Omega0_baryon;
    p.Og = Omega0_gamma;
    p.Ode = 1 - (p.Om + p.Og);
    p.w0 = w0;
    p.wa = wa;
    const double hH0 = h * BG_H0;
    for (int i = 0; i < bins; ++i) {
        const double zi = 15. * i / (double)(bins - 1);
        const double zp = 1 + zi;
        const double w = w_of_z(&p, zi);
        const double wi = wint(&p, zi);
        z[i] = zi;
        a[i] = 1. / zp;
        Hz[i] = sqrt(p.Om * zp * zp * zp + p.Og * zp * zp * zp * zp + p.Ode * wi) * hH0;
        conformal_Hz[i] = a[i] * Hz[i];
        conformal_Hz_prime[i] = -(zp * zp * zp * (2 * zp * p.Og + p.Om) + (1 + 3 * w) * p.Ode * wi) / (zp * zp) / 2. *
                                hH0 * hH0;
    }
    /* comoving distance, cumulative */
    chi[0] = 0.0;
    for (int i = 1; i < bins; ++i) chi[i] = chi[i - 1] + inv_E_integral(&p, z[i - 1], z[i]) / hH0;
    /* growth: integrate from a = 0.05 up through the grid (descending z) */
    double y[2] = {0.05, 1.0};
    double acur = 0.05;
    for (int i = bins - 1; i >= 0; --i) {
        if (a[i] > acur) {
            const int nsub = 64 + (int)(64 * fabs(log(a[i] / acur)) / 0.01);
            growth_step(&p, acur, a[i], y, nsub > 4096 ? 4096 : nsub);
            acur = a[i];
        }
        D1[i] = y[0];
        f[i] = y[1] * a[i] / y[0];
    }
    const double D10 = D1[0];
    for (int i = 0; i < bins; ++i) D1[i] /= D10;
    for (int i = 0; i < bins; ++i) {
        if (z[i] > 1e-10) {
            const double sv1 = lin_interp(bias_z, s1, n_bias, z[i]);
            const double sv2 = lin_interp(bias_z, s2, n_bias, z[i]);
            const double e1 = lin_interp(bias_z, fevo1, n_bias, z[i]);
            const double e2 = lin_interp(bias_z, fevo2, n_bias, z[i]);
            const double base = conformal_Hz_prime[i] / (conformal_Hz[i] * conformal_Hz[i]);
            G1[i] = base + (2 - 5 * sv1) / (chi[i] * conformal_Hz[i]) + 5 * sv1 - e1;
            G2[i] = base + (2 - 5 * sv2) / (chi[i] * conformal_Hz[i]) + 5 * sv2 - e2;
        } else {
            G1[i] = 0;
            G2[i] = 0;
        }
    }
    return 0;
}
