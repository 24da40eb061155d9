/*
 * Oracle (test infrastructure only): multipoles of the local (non-integrated)
 * density, RSD and Doppler (d1) terms, restated from
 *   functions_nonintegrated   functions.c:89-1484 (den-den 129-229, rsd-rsd 230-362,
 *                             d1-d1 363-573, den-rsd 670-771, den-d1 772-855,
 *                             rsd-d1 904-1143; growth factors 1429-1469),
 *   multipoles_nonintegrated_integrand   signal.c:83-111 (mu = 2x - 1, x P_l(mu)),
 *   coffe_integrate NONINTEGRATED/MULTIPOLES   signal.c:521-542 ((2l+1) * QAG),
 *   coffe_multipoles_init     multipoles.c:48-136 (z-major, then l, then r).
 * Background quantities and the I_l^n tables are inputs (Akima splines, as the
 * reference builds them with the default `interp_method`).
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
    size_t n;
    const double *x, *y;
    double *b, *c, *d;
} akima_t;

static int ak_init(akima_t *a, size_t n, const double *x, const double *y)
{
    a->n = n; a->x = x; a->y = y;
    a->b = (double *)malloc(sizeof(double) * 3 * n);
    if (!a->b) return ORC_ENOMEM;
    a->c = a->b + n; a->d = a->c + n;
    return orc_akima_init(n, x, y, a->b, a->c, a->d);
}

static double ak(const akima_t *a, double x)
{
    return orc_akima_eval(a->n, a->x, a->y, a->b, a->c, a->d, x);
}

/* table order: I00, I02, I04, I11, I13, I20, I22, I31 (integrals.c:599-608) */
enum { T00 = 0, T02, T04, T11, T13, T20, T22, T31, NTAB };

typedef struct {
    akima_t tab[NTAB];
    int have[NTAB];
    akima_t chi, z_of_chi, D1, f, cH, G1, G2, b1, b2;
    int den, rsd, d1, flatsky;
    /* d2 and g1-g3 (orc_dg): flags, end-point tables, divergent table and its grid */
    int d2, g1, g2, g3, dg_only_cross;
    double H0h, Omega_m;
    akima_t a, s1, s2, fevo1, fevo2, div;
    orc_bicubic *ren;
} mp_ctx;

typedef struct {
    const mp_ctx *ctx;
    double z_mean, sep;
    int l;
} mp_point;

static double I(const mp_ctx *c, int t, double r)
{
    return ak(&c->tab[t], r);
}

double orc_functions_nonintegrated(const void *vctx, double z_mean, double mu, double sep)
{
    const mp_ctx *c = (const mp_ctx *)vctx;
    const double chi_mean = ak(&c->chi, z_mean);
    const double chi1 = chi_mean - sep * mu / 2.;
    const double chi2 = chi_mean + sep * mu / 2.;
    const double costheta = (2. * chi_mean * chi_mean - sep * sep + mu * mu * sep * sep / 2.) /
                            (2. * chi_mean * chi_mean - mu * mu * sep * sep / 2.);
    const double z1 = ak(&c->z_of_chi, chi1);
    const double z2 = ak(&c->z_of_chi, chi2);
    const double f1 = ak(&c->f, z1), f2 = ak(&c->f, z2), fmean = ak(&c->f, z_mean);
    const double H1 = ak(&c->cH, z1), H2 = ak(&c->cH, z2), Hmean = ak(&c->cH, z_mean);
    const double b1 = ak(&c->b1, z1), b2 = ak(&c->b2, z2);
    const double bm1 = ak(&c->b1, z_mean), bm2 = ak(&c->b2, z_mean);
    const double G1 = ak(&c->G1, z1), G2 = ak(&c->G2, z2);
    const double Gm1 = ak(&c->G1, z_mean), Gm2 = ak(&c->G2, z_mean);
    const double P2 = orc_legendre_pl(2, mu), P4 = orc_legendre_pl(4, mu);
    const double ct2 = costheta * costheta;
    double result = 0;
    if (c->den) {
        if (c->flatsky) result += bm1 * bm2 * I(c, T00, sep);
        else result += b1 * b2 * I(c, T00, sep);
    }
    if (c->rsd) {
        if (c->flatsky) {
            result += fmean * fmean * I(c, T00, sep) / 5. - 4 * fmean * fmean * I(c, T02, sep) / 7. * P2 +
                      8 * fmean * fmean * I(c, T04, sep) / 35. * P4;
        } else {
            result += f1 * f2 * (1 + 2 * ct2) / 15 * I(c, T00, sep) -
                      f1 * f2 / 21. * ((1 + 11. * ct2) + 18 * costheta * (ct2 - 1) * chi1 * chi2 / sep / sep) *
                          I(c, T02, sep) +
                      f1 * f2 *
                          (4 * (3 * ct2 - 1) * (pow(chi1, 4) + pow(chi2, 4)) / 35. / pow(sep, 4) +
                           chi1 * chi2 * (3 + ct2) *
                               (3 * (3 + ct2) * chi1 * chi2 - 8 * (chi1 * chi1 + chi2 * chi2) * costheta) / 35. /
                               pow(sep, 4)) *
                          I(c, T04, sep);
        }
    }
    if (c->d1) {
        if (c->flatsky) {
            result += (Hmean * Hmean * fmean * fmean * Gm1 * Gm2 * I(c, T20, sep) / 3. +
                       2 * Hmean * Hmean * fmean * fmean * Gm1 * Gm2 * pow(sep, 2) * I(c, T22, sep) * P2 / 3.);
        } else {
            result += (H1 * H2 * f1 * f2 * G1 * G2 * costheta * I(c, T20, sep) / 3. -
                       H1 * H2 * f1 * f2 * G1 * G2 *
                           ((chi2 - chi1 * costheta) * (chi1 - chi2 * costheta) + pow(sep, 2) * costheta / 3.) *
                           I(c, T22, sep));
        }
    }
    if (c->den && c->rsd) {
        if (c->flatsky) {
            result += (bm1 * fmean / 3. + bm2 * fmean / 3.) * I(c, T00, sep) -
                      (2 * bm1 * fmean / 3. + 2 * bm2 * fmean / 3.) * I(c, T02, sep) * P2;
        } else {
            result += (b1 * f2 / 3. + b2 * f1 / 3.) * I(c, T00, sep) -
                      (b1 * f2 * (2. / 3. - (1. - ct2) * pow(chi1 / sep, 2)) +
                       b2 * f1 * (2. / 3. - (1. - ct2) * pow(chi2 / sep, 2))) *
                          I(c, T02, sep);
        }
    }
    if (c->den && c->d1) {
        if (c->flatsky) {
            result += (bm1 * fmean * Hmean * Gm2 * mu + bm2 * fmean * Hmean * Gm1 * (-mu)) * sep * I(c, T11, sep);
        } else {
            result += -(b1 * f2 * H2 * G2 * (chi1 * costheta - chi2) + b2 * f1 * H1 * G1 * (chi2 * costheta - chi1)) *
                      I(c, T11, sep);
        }
    }
    if (c->rsd && c->d1) {
        if (c->flatsky) {
            const double P3p = orc_legendre_pl(3, mu), P3m = orc_legendre_pl(3, -mu);
            result += (-3. / 5. * (fmean * fmean * Hmean * Gm2 * (-mu) + fmean * fmean * Hmean * Gm1 * mu) *
                           I(c, T11, sep) * sep +
                       (fmean * fmean * Hmean * Gm2 * 2. / 5. * P3m + fmean * fmean * Hmean * Gm1 * 2. / 5. * P3p) *
                           I(c, T13, sep) * sep);
        } else {
            result += ((f1 * f2 * H2 * G2 * ((1. + 2 * ct2) * chi2 - 3 * chi1 * costheta) / 5. +
                        f2 * f1 * H1 * G1 * ((1. + 2 * ct2) * chi1 - 3 * chi2 * costheta) / 5.) *
                           I(c, T11, sep) +
                       (f1 * f2 * H2 * G2 *
                            ((1. - 3 * ct2) * pow(chi2, 3) + costheta * (5. + ct2) * pow(chi2, 2) * chi1 -
                             2 * (2. + ct2) * chi2 * pow(chi1, 2) + 2 * pow(chi1, 3) * costheta) /
                            5 +
                        f2 * f1 * H1 * G1 *
                            ((1. - 3 * ct2) * pow(chi1, 3) + costheta * (5. + ct2) * pow(chi1, 2) * chi2 -
                             2 * (2. + ct2) * chi1 * pow(chi2, 2) + 2 * pow(chi2, 3) * costheta) /
                            5) *
                           I(c, T13, sep) / pow(sep, 2));
        }
    }
    if (c->d2 || c->g1 || c->g2 || c->g3) {
        /* functions.c:574-669 (auto), 856-903 (den), 1144-1212 (rsd), 1213-1254 (d1),
           1255-1428 (among d2, g1-g3); the end points' values in flat sky too, as there */
        const double a1 = ak(&c->a, z1), a2 = ak(&c->a, z2);
        const double s1 = ak(&c->s1, z1), s2 = ak(&c->s2, z2);
        const double fevo1 = ak(&c->fevo1, z1), fevo2 = ak(&c->fevo2, z2);
        const double Om = c->Omega_m, h2 = pow(c->H0h, 2), h4 = pow(c->H0h, 4);
        const double I40 = ak(&c->div, sep) - orc_bicubic_eval(c->ren, chi1, chi2);
        const double I20 = I(c, T20, sep), I22 = I(c, T22, sep), I31 = I(c, T31, sep);
        const double wr1 = 2. / 3 * pow(sep, 2) - (1 - pow(costheta, 2)) * pow(chi1, 2);
        const double wr2 = 2. / 3 * pow(sep, 2) - (1 - pow(costheta, 2)) * pow(chi2, 2);
        /* the g sources' end-point factors (1 + G), (5 s - 2), (f - 1) */
        const int gon[3] = {c->g1, c->g2, c->g3};
        const double g_1[3] = {1 + G1, 5 * s1 - 2, f1 - 1}, g_2[3] = {1 + G2, 5 * s2 - 2, f2 - 1};
        /* the autos, skipped with only_cross_correlations (functions.c:574-669) */
        if (c->d2 && !c->dg_only_cross) result += (3 - fevo1) * (3 - fevo2) * pow(H1, 2) * pow(H2, 2) * f1 * f2 * I40;
        for (int i = 0; i < 3; ++i)
            if (gon[i] && !c->dg_only_cross) result += h4 * (9 * pow(Om, 2) * g_1[i] * g_2[i] / 4 / a1 / a2 * I40);
        if (c->den && c->d2)
            result += ((3 - fevo2) * b1 * f2 * pow(H2, 2) + (3 - fevo1) * b2 * f1 * pow(H1, 2)) * I20;
        for (int i = 0; i < 3; ++i)
            if (c->den && gon[i])
                result += h2 * (-(b1 * 3 * Om / 2 / a2 * g_2[i] + b2 * 3 * Om / 2 / a1 * g_1[i]) * I20);
        if (c->rsd && c->d2)
            result += ((3 - fevo2) / 3 * f1 * f2 * pow(H2, 2) + (3 - fevo1) / 3 * f2 * f1 * pow(H1, 2)) * I20 -
                      ((3 - fevo2) * f1 * f2 * pow(H2, 2) * wr2 + (3 - fevo1) * f2 * f1 * pow(H1, 2) * wr1) * I22;
        for (int i = 0; i < 3; ++i)
            if (c->rsd && gon[i])
                result += h2 * (-(Om / 2. / a2 * f1 * g_2[i] + Om / 2. / a1 * f2 * g_1[i]) * I20 +
                                (3 * Om / 2. / a2 * f1 * g_2[i] * wr2 + 3 * Om / 2. / a1 * f2 * g_1[i] * wr1) * I22);
        if (c->d1 && c->d2)
            result += -((3 - fevo2) * H1 * pow(H2, 2) * f1 * f2 * (chi2 * costheta - chi1) +
                        (3 - fevo1) * H2 * pow(H1, 2) * f2 * f1 * (chi1 * costheta - chi2)) *
                      I31;
        for (int i = 0; i < 3; ++i)
            if (c->d1 && gon[i])
                result += h2 * ((3 * Om / 2. / a2 * H1 * f1 * g_2[i] * (chi2 * costheta - chi1) +
                                 3 * Om / 2. / a1 * H2 * f2 * g_1[i] * (chi1 * costheta - chi2)) *
                                I31);
        for (int i = 0; i < 3; ++i)
            if (c->d2 && gon[i])
                result += h2 * (-(3 * (3 - fevo1) * Om / 2. / a2 * pow(H1, 2) * f1 * g_2[i] +
                                  3 * (3 - fevo2) * Om / 2. / a1 * pow(H2, 2) * f2 * g_1[i]) *
                                I40);
        for (int i = 0; i < 3; ++i)
            for (int j = i + 1; j < 3; ++j)
                if (gon[i] && gon[j])
                    result += h4 * (9 * pow(Om, 2) / 4. * (g_1[i] * g_2[j] + g_2[i] * g_1[j]) / a1 / a2 * I40);
    }
    if (!isfinite(result)) return NAN;
    if (c->flatsky) return result * pow(ak(&c->D1, z_mean), 2);
    return result * ak(&c->D1, z1) * ak(&c->D1, z2);
}

static double mp_integrand(double x, void *vp)
{
    const mp_point *p = (const mp_point *)vp;
    const double mu = 2 * x - 1;
    const double v = orc_functions_nonintegrated(p->ctx, p->z_mean, mu, p->sep);
    return p->l == 0 ? v : v * orc_legendre_pl(p->l, mu);
}

/*
 * tables: NTAB entries (x, y, n); an entry with n == 0 is absent.
 * bg: z grid (nbg) with chi, D1, f, conformal H, G1, G2; bias tables on zb (nb).
 * quad: 0 = QAG-GK61 with epsrel, M > 0 = M-point Gauss-Legendre in x.
 */
/* the d2 / g1-g3 part of the context (dg may be NULL: none) */
static int dg_init(mp_ctx *c, size_t nbg, const double *bg_z, const orc_dg *dg)
{
    if (!dg || !(dg->d2 || dg->g1 || dg->g2 || dg->g3)) return ORC_OK;
    c->d2 = dg->d2; c->g1 = dg->g1; c->g2 = dg->g2; c->g3 = dg->g3;
    c->dg_only_cross = dg->only_cross;
    c->H0h = dg->H0h;
    c->Omega_m = dg->Omega_m;
    int st = ak_init(&c->a, nbg, bg_z, dg->bg_a);
    if (st == ORC_OK) st = ak_init(&c->s1, dg->nmb, dg->mb_z, dg->s1);
    if (st == ORC_OK) st = ak_init(&c->s2, dg->nmb, dg->mb_z, dg->s2);
    if (st == ORC_OK) st = ak_init(&c->fevo1, dg->nmb, dg->mb_z, dg->fevo1);
    if (st == ORC_OK) st = ak_init(&c->fevo2, dg->nmb, dg->mb_z, dg->fevo2);
    if (st == ORC_OK) st = ak_init(&c->div, dg->ndiv, dg->div_x, dg->div_y);
    if (st == ORC_OK) {
        c->ren = orc_bicubic_new(dg->nren, dg->nren, dg->ren_chi, dg->ren_chi, dg->ren);
        if (!c->ren) st = ORC_ENOMEM;
    }
    return st;
}

static void dg_free(mp_ctx *c)
{
    free(c->a.b); free(c->s1.b); free(c->s2.b); free(c->fevo1.b); free(c->fevo2.b); free(c->div.b);
    orc_bicubic_free(c->ren);
}

int orc_multipoles(const double *const *tab_x, const double *const *tab_y, const size_t *tab_n, size_t nbg,
                   const double *bg_z, const double *bg_chi, const double *bg_D1, const double *bg_f,
                   const double *bg_cH, const double *bg_G1, const double *bg_G2, size_t nb, const double *zb,
                   const double *bias1, const double *bias2, int den, int rsd, int d1, int flatsky, size_t nz,
                   const double *z_mean, size_t nl, const int *ls, size_t nsep, const double *sep, int quad,
                   double epsrel, double *out)
{
    return orc_multipoles_dg(tab_x, tab_y, tab_n, nbg, bg_z, bg_chi, bg_D1, bg_f, bg_cH, bg_G1, bg_G2, nb, zb, bias1,
                             bias2, den, rsd, d1, flatsky, nz, z_mean, nl, ls, nsep, sep, quad, epsrel, NULL, out);
}

static int mp_ctx_init(mp_ctx *cp, const double *const *tab_x, const double *const *tab_y, const size_t *tab_n,
                       size_t nbg, const double *bg_z, const double *bg_chi, const double *bg_D1, const double *bg_f,
                       const double *bg_cH, const double *bg_G1, const double *bg_G2, size_t nb, const double *zb,
                       const double *bias1, const double *bias2, int den, int rsd, int d1, int flatsky,
                       const orc_dg *dg)
{
    mp_ctx c;
    memset(&c, 0, sizeof(c));
    c.den = den; c.rsd = rsd; c.d1 = d1; c.flatsky = flatsky;
    int st = ORC_OK;
    for (int t = 0; t < NTAB && st == ORC_OK; ++t) {
        if (tab_n[t]) { c.have[t] = 1; st = ak_init(&c.tab[t], tab_n[t], tab_x[t], tab_y[t]); }
    }
    if (st == ORC_OK) st = ak_init(&c.chi, nbg, bg_z, bg_chi);
    if (st == ORC_OK) st = ak_init(&c.z_of_chi, nbg, bg_chi, bg_z);
    if (st == ORC_OK) st = ak_init(&c.D1, nbg, bg_z, bg_D1);
    if (st == ORC_OK) st = ak_init(&c.f, nbg, bg_z, bg_f);
    if (st == ORC_OK) st = ak_init(&c.cH, nbg, bg_z, bg_cH);
    if (st == ORC_OK) st = ak_init(&c.G1, nbg, bg_z, bg_G1);
    if (st == ORC_OK) st = ak_init(&c.G2, nbg, bg_z, bg_G2);
    if (st == ORC_OK) st = ak_init(&c.b1, nb, zb, bias1);
    if (st == ORC_OK) st = ak_init(&c.b2, nb, zb, bias2);
    if (st == ORC_OK) st = dg_init(&c, nbg, bg_z, dg);
    *cp = c;
    return st;
}

static void mp_ctx_free(mp_ctx *c)
{
    for (int t = 0; t < NTAB; ++t) if (c->have[t]) free(c->tab[t].b);
    free(c->chi.b); free(c->z_of_chi.b); free(c->D1.b); free(c->f.b); free(c->cH.b); free(c->G1.b); free(c->G2.b);
    free(c->b1.b); free(c->b2.b);
    dg_free(c);
}

int orc_multipoles_dg(const double *const *tab_x, const double *const *tab_y, const size_t *tab_n, size_t nbg,
                      const double *bg_z, const double *bg_chi, const double *bg_D1, const double *bg_f,
                      const double *bg_cH, const double *bg_G1, const double *bg_G2, size_t nb, const double *zb,
                      const double *bias1, const double *bias2, int den, int rsd, int d1, int flatsky, size_t nz,
                      const double *z_mean, size_t nl, const int *ls, size_t nsep, const double *sep, int quad,
                      double epsrel, const orc_dg *dg, double *out)
{
    mp_ctx c;
    int st = mp_ctx_init(&c, tab_x, tab_y, tab_n, nbg, bg_z, bg_chi, bg_D1, bg_f, bg_cH, bg_G1, bg_G2, nb, zb, bias1,
                         bias2, den, rsd, d1, flatsky, dg);
    double *glx = NULL, *glw = NULL;
    if (st == ORC_OK && quad > 0) {
        glx = (double *)malloc(sizeof(double) * (size_t)quad);
        glw = (double *)malloc(sizeof(double) * (size_t)quad);
        orc_gauss_legendre((size_t)quad, glx, glw);
    }
    const size_t total = nz * nl * nsep;
    if (st == ORC_OK) {
        #pragma omp parallel for schedule(dynamic)
        for (size_t idx = 0; idx < total; ++idx) {
            const size_t iz = idx / (nl * nsep), il = (idx / nsep) % nl, ir = idx % nsep;
            mp_point p = {&c, z_mean[iz], sep[ir], ls[il]};
            double v;
            if (quad > 0) {
                double s = 0.0;
                for (int q = 0; q < quad; ++q) s += glw[q] * mp_integrand(0.5 * (glx[q] + 1.0), &p);
                v = 0.5 * s;
            } else {
                v = orc_qag(mp_integrand, &p, 0, 1, 0, epsrel, 50000, NULL, NULL);
            }
            out[idx] = (2 * ls[il] + 1) * v;
        }
    }
    free(glx); free(glw);
    mp_ctx_free(&c);
    return st;
}

/*
 * The local correlation function (functions_nonintegrated, functions.c:89-1484, the
 * reference's CORRFUNC output of signal.c:504-509) at every (z_mean, mu, sep):
 * out[z][mu][r].  Same inputs as orc_multipoles_dg.
 */
int orc_corrfunc_dg(const double *const *tab_x, const double *const *tab_y, const size_t *tab_n, size_t nbg,
                    const double *bg_z, const double *bg_chi, const double *bg_D1, const double *bg_f,
                    const double *bg_cH, const double *bg_G1, const double *bg_G2, size_t nb, const double *zb,
                    const double *bias1, const double *bias2, int den, int rsd, int d1, int flatsky, size_t nz,
                    const double *z_mean, size_t nmu, const double *mu, size_t nsep, const double *sep,
                    const orc_dg *dg, double *out)
{
    mp_ctx c;
    int st = mp_ctx_init(&c, tab_x, tab_y, tab_n, nbg, bg_z, bg_chi, bg_D1, bg_f, bg_cH, bg_G1, bg_G2, nb, zb, bias1,
                         bias2, den, rsd, d1, flatsky, dg);
    if (st == ORC_OK)
        for (size_t iz = 0; iz < nz; ++iz)
            for (size_t im = 0; im < nmu; ++im)
                for (size_t ir = 0; ir < nsep; ++ir)
                    out[(iz * nmu + im) * nsep + ir] = orc_functions_nonintegrated(&c, z_mean[iz], mu[im], sep[ir]);
    mp_ctx_free(&c);
    return st;
}

/*
 * CPU baseline: the reference's own sequence for each spectrum -- the 8 default
 * tables of coffe_integrals_init (integrals.c:599-660, with or without the 43
 * small-r QAG nodes) followed by the QAG mu-integral of every (z, l, r) point
 * (multipoles.c:88-99, signal.c:521-542) -- with OpenMP over spectra.
 * out: [spec][z][l][r].
 */
int orc_reference_batch(int nspec, const long long *off, const double *k, const double *p, const long long *nfft,
                        size_t nbg, const double *bg_z, const double *bg_chi, const double *bg_D1, const double *bg_f,
                        const double *bg_cH, const double *bg_G1, const double *bg_G2, size_t nb, const double *zb,
                        const double *bias1, const double *bias2, int den, int rsd, int d1, int flatsky, size_t nz,
                        const double *z_mean, size_t nl, const int *ls, size_t nsep, const double *sep, int small_r,
                        double epsrel, double *out)
{
    int status = ORC_OK;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int s = 0; s < nspec; ++s) {
        const size_t nin = (size_t)(off[s + 1] - off[s]);
        const size_t N = (size_t)nfft[s];
        const double *ks = k + off[s], *ps = p + off[s];
        double *tx[NTAB], *ty[NTAB];
        size_t tn[NTAB];
        int st = ORC_OK;
        static const int tabs[NTAB][2] = {{0, 0}, {0, 2}, {0, 4}, {1, 1}, {1, 3}, {2, 0}, {2, 2}, {3, 1}};
        for (int t = 0; t < NTAB; ++t) {
            tx[t] = (double *)malloc(sizeof(double) * (N + 43));
            ty[t] = (double *)malloc(sizeof(double) * (N + 43));
            tn[t] = N + 43;
            if (st == ORC_OK)
                st = orc_integral_table(tabs[t][0], tabs[t][1], N, nin, ks, ps, ks[0], ks[nin - 1], small_r, tx[t],
                                        ty[t]);
        }
        mp_ctx c;
        memset(&c, 0, sizeof(c));
        c.den = den; c.rsd = rsd; c.d1 = d1; c.flatsky = flatsky;
        for (int t = 0; t < NTAB && st == ORC_OK; ++t) { c.have[t] = 1; st = ak_init(&c.tab[t], tn[t], tx[t], ty[t]); }
        if (st == ORC_OK) st = ak_init(&c.chi, nbg, bg_z, bg_chi);
        if (st == ORC_OK) st = ak_init(&c.z_of_chi, nbg, bg_chi, bg_z);
        if (st == ORC_OK) st = ak_init(&c.D1, nbg, bg_z, bg_D1);
        if (st == ORC_OK) st = ak_init(&c.f, nbg, bg_z, bg_f);
        if (st == ORC_OK) st = ak_init(&c.cH, nbg, bg_z, bg_cH);
        if (st == ORC_OK) st = ak_init(&c.G1, nbg, bg_z, bg_G1);
        if (st == ORC_OK) st = ak_init(&c.G2, nbg, bg_z, bg_G2);
        if (st == ORC_OK) st = ak_init(&c.b1, nb, zb, bias1);
        if (st == ORC_OK) st = ak_init(&c.b2, nb, zb, bias2);
        for (size_t idx = 0; idx < nz * nl * nsep && st == ORC_OK; ++idx) {
            const size_t iz = idx / (nl * nsep), il = (idx / nsep) % nl, ir = idx % nsep;
            mp_point pt = {&c, z_mean[iz], sep[ir], ls[il]};
            const double v = orc_qag(mp_integrand, &pt, 0, 1, 0, epsrel, 50000, NULL, NULL);
            out[(size_t)s * nz * nl * nsep + idx] = (2 * ls[il] + 1) * v;
        }
        for (int t = 0; t < NTAB; ++t) { if (c.have[t]) free(c.tab[t].b); free(tx[t]); free(ty[t]); }
        free(c.chi.b); free(c.z_of_chi.b); free(c.D1.b); free(c.f.b); free(c.cH.b); free(c.G1.b); free(c.G2.b);
        free(c.b1.b); free(c.b2.b);
        if (st != ORC_OK) {
            #pragma omp atomic write
            status = st;
        }
    }
    return status;
}

/* ---- high-precision weights ---------------------------------------------------- */
/*
 * The mu-integrated weights of the local terms with the reference's formulas
 * (functions.c:129-1143, full and flat sky, written as in the reference: polynomials in
 * cos theta with their (chi/s)^n factors) evaluated in binary128 (__float128, 113-bit
 * significand), so that the cancellations of those terms at s << chi (up to (chi/s)^4 ~
 * 1e10) leave ~1e-24 of error instead of FP64's ~1e-6.  The background and bias values
 * at the end points are the FP64 Akima values of the tables (smooth, no cancellation),
 * then carried in binary128.  For the Gauss-Legendre rule (mu_q, w_q) on [-1, 1]:
 *   A[z][l][t][r] = (2l+1)/2 sum_q w_q P_l(mu_q) G(mu_q) c_t(mu_q),
 * t over the 8 standard tables (I00, I02, I04, I11, I13, I20, I22, I31), G = D1(z1)
 * D1(z2) (full sky) or D1(z_mean)^2 (flat sky), functions.c:1457-1467.  terms: the
 * den/rsd/d1 auto and cross terms as flags (den, rsd, d1 as in orc_multipoles).
 */
typedef __float128 qf;

static qf qpl(int l, qf x)
{
    if (l == 0) return 1;
    qf p0 = 1, p1 = x;
    for (int k = 2; k <= l; ++k) {
        const qf p2 = ((qf)(2 * k - 1) * x * p1 - (qf)(k - 1) * p0) / (qf)k;
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

/* out[0..NTAB) the standard tables' coefficients, out[NTAB] that of I_0^4 and
   out[NTAB + 1] minus that of the renormalisation grid's value at (chi1, chi2) (the d2 /
   g terms), all times G */
static void hp_coefficients(const mp_ctx *c, double z_mean, double mu_d, double sep_d, qf *out)
{
    for (int t = 0; t < NTAB + 2; ++t) out[t] = 0;
    const double chi_mean_d = ak(&c->chi, z_mean);
    const qf chi_mean = chi_mean_d, mu = mu_d, sep = sep_d;
    const qf chi1 = chi_mean - sep * mu / 2;
    const qf chi2 = chi_mean + sep * mu / 2;
    const qf costheta = (2 * chi_mean * chi_mean - sep * sep + mu * mu * sep * sep / 2) /
                        (2 * chi_mean * chi_mean - mu * mu * sep * sep / 2);
    /* the end points' redshifts from the FP64 chi1, chi2 the reference computes */
    const double chi1_d = chi_mean_d - sep_d * mu_d / 2., chi2_d = chi_mean_d + sep_d * mu_d / 2.;
    const double z1 = ak(&c->z_of_chi, chi1_d), z2 = ak(&c->z_of_chi, chi2_d);
    const qf f1 = ak(&c->f, z1), f2 = ak(&c->f, z2), fmean = ak(&c->f, z_mean);
    const qf H1 = ak(&c->cH, z1), H2 = ak(&c->cH, z2), Hmean = ak(&c->cH, z_mean);
    const qf b1 = ak(&c->b1, z1), b2 = ak(&c->b2, z2);
    const qf bm1 = ak(&c->b1, z_mean), bm2 = ak(&c->b2, z_mean);
    const qf G1 = ak(&c->G1, z1), G2 = ak(&c->G2, z2);
    const qf Gm1 = ak(&c->G1, z_mean), Gm2 = ak(&c->G2, z_mean);
    const qf P2 = qpl(2, mu), P4 = qpl(4, mu);
    const qf ct2 = costheta * costheta;
    const qf s2 = sep * sep, s4 = s2 * s2;
    if (c->den) {
        if (c->flatsky) out[T00] += bm1 * bm2;
        else out[T00] += b1 * b2;
    }
    if (c->rsd) {
        if (c->flatsky) {
            out[T00] += fmean * fmean / 5;
            out[T02] += -4 * fmean * fmean / 7 * P2;
            out[T04] += 8 * fmean * fmean / 35 * P4;
        } else {
            out[T00] += f1 * f2 * (1 + 2 * ct2) / 15;
            out[T02] += -f1 * f2 / 21 * ((1 + 11 * ct2) + 18 * costheta * (ct2 - 1) * chi1 * chi2 / sep / sep);
            out[T04] += f1 * f2 *
                        (4 * (3 * ct2 - 1) * (chi1 * chi1 * chi1 * chi1 + chi2 * chi2 * chi2 * chi2) / 35 / s4 +
                         chi1 * chi2 * (3 + ct2) * (3 * (3 + ct2) * chi1 * chi2 - 8 * (chi1 * chi1 + chi2 * chi2) * costheta) /
                             35 / s4);
        }
    }
    if (c->d1) {
        if (c->flatsky) {
            out[T20] += Hmean * Hmean * fmean * fmean * Gm1 * Gm2 / 3;
            out[T22] += 2 * Hmean * Hmean * fmean * fmean * Gm1 * Gm2 * s2 * P2 / 3;
        } else {
            out[T20] += H1 * H2 * f1 * f2 * G1 * G2 * costheta / 3;
            out[T22] += -H1 * H2 * f1 * f2 * G1 * G2 *
                        ((chi2 - chi1 * costheta) * (chi1 - chi2 * costheta) + s2 * costheta / 3);
        }
    }
    if (c->den && c->rsd) {
        if (c->flatsky) {
            out[T00] += bm1 * fmean / 3 + bm2 * fmean / 3;
            out[T02] += -(2 * bm1 * fmean / 3 + 2 * bm2 * fmean / 3) * P2;
        } else {
            out[T00] += b1 * f2 / 3 + b2 * f1 / 3;
            out[T02] += -(b1 * f2 * ((qf)2 / 3 - (1 - ct2) * (chi1 / sep) * (chi1 / sep)) +
                          b2 * f1 * ((qf)2 / 3 - (1 - ct2) * (chi2 / sep) * (chi2 / sep)));
        }
    }
    if (c->den && c->d1) {
        if (c->flatsky) {
            out[T11] += (bm1 * fmean * Hmean * Gm2 * mu + bm2 * fmean * Hmean * Gm1 * (-mu)) * sep;
        } else {
            out[T11] += -(b1 * f2 * H2 * G2 * (chi1 * costheta - chi2) + b2 * f1 * H1 * G1 * (chi2 * costheta - chi1));
        }
    }
    if (c->rsd && c->d1) {
        if (c->flatsky) {
            const qf P3p = qpl(3, mu), P3m = qpl(3, -mu);
            out[T11] += -(qf)3 / 5 * (fmean * fmean * Hmean * Gm2 * (-mu) + fmean * fmean * Hmean * Gm1 * mu) * sep;
            out[T13] += (fmean * fmean * Hmean * Gm2 * 2 / 5 * P3m + fmean * fmean * Hmean * Gm1 * 2 / 5 * P3p) * sep;
        } else {
            out[T11] += f1 * f2 * H2 * G2 * ((1 + 2 * ct2) * chi2 - 3 * chi1 * costheta) / 5 +
                        f2 * f1 * H1 * G1 * ((1 + 2 * ct2) * chi1 - 3 * chi2 * costheta) / 5;
            out[T13] += (f1 * f2 * H2 * G2 *
                             ((1 - 3 * ct2) * chi2 * chi2 * chi2 + costheta * (5 + ct2) * chi2 * chi2 * chi1 -
                              2 * (2 + ct2) * chi2 * chi1 * chi1 + 2 * chi1 * chi1 * chi1 * costheta) /
                             5 +
                         f2 * f1 * H1 * G1 *
                             ((1 - 3 * ct2) * chi1 * chi1 * chi1 + costheta * (5 + ct2) * chi1 * chi1 * chi2 -
                              2 * (2 + ct2) * chi1 * chi2 * chi2 + 2 * chi2 * chi2 * chi2 * costheta) /
                             5) /
                        s2;
        }
    }
    if (c->d2 || c->g1 || c->g2 || c->g3) {
        /* functions.c:574-669, 856-903, 1144-1428 as in orc_functions_nonintegrated */
        const qf a1 = ak(&c->a, z1), a2 = ak(&c->a, z2);
        const qf mb1 = ak(&c->s1, z1), mb2 = ak(&c->s2, z2);   /* magnification biases */
        const qf fevo1 = ak(&c->fevo1, z1), fevo2 = ak(&c->fevo2, z2);
        const qf Om = c->Omega_m, h2 = (qf)c->H0h * (qf)c->H0h, h4 = h2 * h2;
        const qf ren = orc_bicubic_eval(c->ren, chi1_d, chi2_d);
        const qf wr1 = (qf)2 / 3 * s2 - (1 - ct2) * chi1 * chi1;   /* s2 = sep^2 */
        const qf wr2 = (qf)2 / 3 * s2 - (1 - ct2) * chi2 * chi2;
        const int gon[3] = {c->g1, c->g2, c->g3};
        const qf g_1[3] = {1 + G1, 5 * mb1 - 2, f1 - 1}, g_2[3] = {1 + G2, 5 * mb2 - 2, f2 - 1};
        qf k40 = 0;   /* the coefficient of I_0^4 - ren */
        if (c->d2) k40 += (3 - fevo1) * (3 - fevo2) * H1 * H1 * H2 * H2 * f1 * f2;
        for (int i = 0; i < 3; ++i)
            if (gon[i]) k40 += h4 * (9 * Om * Om * g_1[i] * g_2[i] / 4 / a1 / a2);
        for (int i = 0; i < 3; ++i)
            if (c->d2 && gon[i])
                k40 += h2 * (-(3 * (3 - fevo1) * Om / 2 / a2 * H1 * H1 * f1 * g_2[i] +
                               3 * (3 - fevo2) * Om / 2 / a1 * H2 * H2 * f2 * g_1[i]));
        for (int i = 0; i < 3; ++i)
            for (int j = i + 1; j < 3; ++j)
                if (gon[i] && gon[j]) k40 += h4 * (9 * Om * Om / 4 * (g_1[i] * g_2[j] + g_2[i] * g_1[j]) / a1 / a2);
        out[NTAB] += k40;
        out[NTAB + 1] -= k40 * ren;
        if (c->den && c->d2) out[T20] += (3 - fevo2) * b1 * f2 * H2 * H2 + (3 - fevo1) * b2 * f1 * H1 * H1;
        for (int i = 0; i < 3; ++i)
            if (c->den && gon[i]) out[T20] += h2 * (-(b1 * 3 * Om / 2 / a2 * g_2[i] + b2 * 3 * Om / 2 / a1 * g_1[i]));
        if (c->rsd && c->d2) {
            out[T20] += (3 - fevo2) / 3 * f1 * f2 * H2 * H2 + (3 - fevo1) / 3 * f2 * f1 * H1 * H1;
            out[T22] -= (3 - fevo2) * f1 * f2 * H2 * H2 * wr2 + (3 - fevo1) * f2 * f1 * H1 * H1 * wr1;
        }
        for (int i = 0; i < 3; ++i)
            if (c->rsd && gon[i]) {
                out[T20] += h2 * (-(Om / 2 / a2 * f1 * g_2[i] + Om / 2 / a1 * f2 * g_1[i]));
                out[T22] += h2 * (3 * Om / 2 / a2 * f1 * g_2[i] * wr2 + 3 * Om / 2 / a1 * f2 * g_1[i] * wr1);
            }
        if (c->d1 && c->d2)
            out[T31] += -((3 - fevo2) * H1 * H2 * H2 * f1 * f2 * (chi2 * costheta - chi1) +
                          (3 - fevo1) * H2 * H1 * H1 * f2 * f1 * (chi1 * costheta - chi2));
        for (int i = 0; i < 3; ++i)
            if (c->d1 && gon[i])
                out[T31] += h2 * (3 * Om / 2 / a2 * H1 * f1 * g_2[i] * (chi2 * costheta - chi1) +
                                  3 * Om / 2 / a1 * H2 * f2 * g_1[i] * (chi1 * costheta - chi2));
    }
    const qf G = c->flatsky ? (qf)ak(&c->D1, z_mean) * (qf)ak(&c->D1, z_mean) : (qf)ak(&c->D1, z1) * (qf)ak(&c->D1, z2);
    for (int t = 0; t < NTAB + 2; ++t) out[t] *= G;
}

int orc_mu_weights_hp(size_t nbg, const double *bg_z, const double *bg_chi, const double *bg_D1, const double *bg_f,
                      const double *bg_cH, const double *bg_G1, const double *bg_G2, size_t nb, const double *zb,
                      const double *bias1, const double *bias2, int den, int rsd, int d1, int flatsky, size_t nz,
                      const double *z_mean, size_t nl, const int *ls, size_t nsep, const double *sep, size_t nmu,
                      const double *mu, const double *wmu, double *out)
{
    return orc_mu_weights_hp_dg(nbg, bg_z, bg_chi, bg_D1, bg_f, bg_cH, bg_G1, bg_G2, nb, zb, bias1, bias2, den, rsd,
                                d1, flatsky, nz, z_mean, nl, ls, nsep, sep, nmu, mu, wmu, NULL, out, NULL);
}

/* with the d2 / g terms of dg (NULL: none): out [nz][nl][8 or 9][nsep], the ninth table
   I_0^4 when dg is given, and out_ren [nz][nl][nsep] (with dg) = the renormalisation
   grid's part of the multipoles */
int orc_mu_weights_hp_dg(size_t nbg, const double *bg_z, const double *bg_chi, const double *bg_D1,
                         const double *bg_f, const double *bg_cH, const double *bg_G1, const double *bg_G2, size_t nb,
                         const double *zb, const double *bias1, const double *bias2, int den, int rsd, int d1,
                         int flatsky, size_t nz, const double *z_mean, size_t nl, const int *ls, size_t nsep,
                         const double *sep, size_t nmu, const double *mu, const double *wmu, const orc_dg *dg,
                         double *out, double *out_ren)
{
    mp_ctx c;
    memset(&c, 0, sizeof(c));
    c.den = den; c.rsd = rsd; c.d1 = d1; c.flatsky = flatsky;
    int st = ak_init(&c.chi, nbg, bg_z, bg_chi);
    if (st == ORC_OK) st = ak_init(&c.z_of_chi, nbg, bg_chi, bg_z);
    if (st == ORC_OK) st = ak_init(&c.D1, nbg, bg_z, bg_D1);
    if (st == ORC_OK) st = ak_init(&c.f, nbg, bg_z, bg_f);
    if (st == ORC_OK) st = ak_init(&c.cH, nbg, bg_z, bg_cH);
    if (st == ORC_OK) st = ak_init(&c.G1, nbg, bg_z, bg_G1);
    if (st == ORC_OK) st = ak_init(&c.G2, nbg, bg_z, bg_G2);
    if (st == ORC_OK) st = ak_init(&c.b1, nb, zb, bias1);
    if (st == ORC_OK) st = ak_init(&c.b2, nb, zb, bias2);
    if (st == ORC_OK) st = dg_init(&c, nbg, bg_z, dg);
    const int with_dg = c.d2 || c.g1 || c.g2 || c.g3;
    const size_t nt = with_dg ? NTAB + 1 : NTAB;
    if (st == ORC_OK) {
        #pragma omp parallel for schedule(dynamic)
        for (size_t idx = 0; idx < nz * nsep; ++idx) {
            const size_t iz = idx / nsep, ir = idx % nsep;
            qf acc[4][NTAB + 2];
            memset(acc, 0, sizeof(acc));
            for (size_t q = 0; q < nmu; ++q) {
                qf cq[NTAB + 2];
                hp_coefficients(&c, z_mean[iz], mu[q], sep[ir], cq);
                for (size_t il = 0; il < nl && il < 4; ++il) {
                    const qf pw = (qf)wmu[q] * qpl(ls[il], (qf)mu[q]);
                    for (int t = 0; t < NTAB + 2; ++t) acc[il][t] += pw * cq[t];
                }
            }
            for (size_t il = 0; il < nl && il < 4; ++il) {
                const qf norm = (qf)(2 * ls[il] + 1) / 2;
                for (size_t t = 0; t < nt; ++t) out[((iz * nl + il) * nt + t) * nsep + ir] = (double)(norm * acc[il][t]);
                if (with_dg && out_ren) out_ren[(iz * nl + il) * nsep + ir] = (double)(norm * acc[il][NTAB + 1]);
            }
        }
    }
    free(c.chi.b); free(c.z_of_chi.b); free(c.D1.b); free(c.f.b); free(c.cH.b); free(c.G1.b); free(c.G2.b);
    free(c.b1.b); free(c.b2.b);
    dg_free(&c);
    return st;
}

/* OpenMP threads of the oracle's parallel loops (the CPU baseline is timed at the box's
   thread count and at one thread); returns the previous setting */
int orc_set_threads(int n)
{
    const int prev = omp_get_max_threads();
    if (n > 0) omp_set_num_threads(n);
    return prev;
}
