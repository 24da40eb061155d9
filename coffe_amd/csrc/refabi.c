/*
 * The reference's C entry points of the table, multipole and covariance paths
 * (integrals.c, signal.c, multipoles.c, average_multipoles.c, covariance.c) over the
 * engine.  The heavy parts run on the GPU: the FFTLog tables (coffe_integral_table), the
 * divergent table and its renormalisations (coffe_divergent_tables), the mu quadrature
 * of the local terms (coffe_mu_weights_ex / _rule) and the covariance integrals and
 * assembly (coffe_cov_integrals / coffe_cov_assemble).  The reference's interpolation
 * objects are made, evaluated and freed by the host program's own routines (common.c),
 * so that every table the reference's callers read back is theirs in type and method.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "ce_host.h"
#include "coffe_refabi.h"
#include "coffe_twobessel.h"
#include "gk61_nodes.h"

#define REF_H0 (1.0 / 2997.92458)   /* COFFE_H0 (common.h:75-77), H0 in h/Mpc */
#define REF_MU_NODES 256            /* the mu rule of the multipoles */
#define REF_Z_NODES 32              /* the redshift rule of the averaged multipoles */
#define REF_REN_BINS 200            /* the renormalisation grid, integrals.c:841 */
#define REF_MAX_L 4                 /* multipoles per weights launch */

/* ---- the host's interpolation routines -------------------------------------------- */

/* common.c's routines, weak: non-NULL when the host program defines them */
extern int coffe_init_spline(coffe_interpolation *, const double *, const double *, size_t, int)
    __attribute__((weak));
extern int coffe_init_spline2d(coffe_interpolation2d *, const double *, const double *, const double *, size_t,
                               size_t, int) __attribute__((weak));
extern double coffe_interp_spline(const coffe_interpolation *, double) __attribute__((weak));
extern double coffe_interp_spline2d(const coffe_interpolation2d *, double, double) __attribute__((weak));
extern int coffe_free_spline(coffe_interpolation *) __attribute__((weak));
extern int coffe_free_spline2d(coffe_interpolation2d *) __attribute__((weak));

static coffe_spline_ops g_ops;
static int g_ops_set;

void coffe_refabi_set_spline_ops(const coffe_spline_ops *ops)
{
    if (ops) {
        g_ops = *ops;
        g_ops_set = 1;
    } else {
        memset(&g_ops, 0, sizeof(g_ops));
        g_ops_set = 0;
    }
}

static int spline_ops(coffe_spline_ops *o)
{
    if (g_ops_set) {
        *o = g_ops;
    } else {
        o->init_spline = coffe_init_spline;
        o->init_spline2d = coffe_init_spline2d;
        o->interp_spline = coffe_interp_spline;
        o->interp_spline2d = coffe_interp_spline2d;
        o->free_spline = coffe_free_spline;
        o->free_spline2d = coffe_free_spline2d;
    }
    return (o->init_spline && o->init_spline2d && o->interp_spline && o->free_spline && o->free_spline2d)
               ? COFFE_OK
               : COFFE_EUNSUPPORTED;
}

/* GSL's gsl_spline (interp, x, y, size) and gsl_spline2d (the gsl_interp2d object, then
   xarr, yarr, zarr; gsl_interp2d: type, xmin, xmax, ymin, ymax, xsize, ysize, state) */
typedef struct {
    const void *interp;
    double *x, *y;
    size_t size;
} gsl_spline_view;
typedef struct {
    const void *type;
    double xmin, xmax, ymin, ymax;
    size_t xsize, ysize;
    void *state;
} gsl_interp2d_view;
typedef struct {
    gsl_interp2d_view obj;
    double *xarr, *yarr, *zarr;
} gsl_spline2d_view;

typedef struct {
    const double *x, *y;
    size_t n;
} nodes;

static int nodes_of(const coffe_interpolation *ip, nodes *out)
{
    const gsl_spline_view *s = ip ? (const gsl_spline_view *)ip->spline : NULL;
    if (!s || !s->x || !s->y || s->size < 2) return COFFE_EINVAL;
    out->x = s->x;
    out->y = s->y;
    out->n = s->size;
    return COFFE_OK;
}

static int same_nodes(const nodes *a, const nodes *b)
{
    return a->n == b->n && (a->x == b->x || memcmp(a->x, b->x, sizeof(double) * a->n) == 0);
}

static int done(int code)
{
    ce_set_last_error(code);
    return code;
}

/* ---- integrals.c ------------------------------------------------------------------ */

static const int STD_N[8] = {0, 0, 0, 1, 1, 2, 2, 3};
/* the scale of the (1/2, -1/2) table (integrals.c:956-962) */
#define REF_FLAT_TSCALE (sqrt(2.0 / M_PI) * 2.0 * M_PI * M_PI)
static const int STD_L[8] = {0, 2, 4, 1, 3, 0, 2, 1};

coffe_integral_t *coffe_find_integral(const coffe_integral_array_t *integral, int n, int l, int state_n,
                                      int state_l)
{
    if (!integral || !integral->array) return NULL;
    for (size_t i = 0; i < integral->size; ++i) {
        coffe_integral_t *t = &integral->array[i];
        if (t->n == n && t->l == l && t->state_n == state_n && t->state_l == state_l) return t;
    }
    return NULL;
}

int coffe_integrals_free(coffe_integral_array_t *integral)
{
    if (!integral) return done(COFFE_EINVAL);
    if (integral->size && integral->array) {
        coffe_spline_ops ops;
        const int st = spline_ops(&ops);
        if (st != COFFE_OK) return done(st);
        for (size_t i = 0; i < integral->size; ++i) {
            coffe_integral_t *t = &integral->array[i];
            ops.free_spline(&t->result);
            ops.free_spline(&t->renormalization_zero_separation);
            ops.free_spline2d(&t->renormalization);
            ops.free_spline2d(&t->result2d);
        }
    }
    free(integral->array);
    integral->array = NULL;
    integral->size = 0;
    return done(COFFE_OK);
}

/* the separations of the divergent table I_0^4 (integrals.c:76-96): a few small ones,
   then runs of constant step */
static size_t divergent_separations(double *out)
{
    static const double head[] = {1e-6, 5e-6, 1e-4, 5e-3, 0.01, 0.02, 0.05, 0.1, 0.2, 0.4, 0.8, 1.0, 1.5};
    /* {first, last, step} */
    static const double runs[][3] = {{2, 30, 1},     {32, 80, 2},      {81, 95, 1},      {95.5, 105, 0.5},
                                     {106, 110, 1},  {112, 120, 2},    {124, 180, 4},    {185, 240, 5},
                                     {250, 300, 10}, {320, 700, 20},   {750, 6600, 50},  {7000, 20000, 1000}};
    size_t n = 0;
    for (size_t i = 0; i < sizeof(head) / sizeof(head[0]); ++i) {
        if (out) out[n] = head[i];
        ++n;
    }
    for (size_t r = 0; r < sizeof(runs) / sizeof(runs[0]); ++r) {
        const int cnt = (int)lround((runs[r][1] - runs[r][0]) / runs[r][2]) + 1;
        for (int j = 0; j < cnt; ++j) {
            if (out) out[n] = runs[r][0] + j * runs[r][2];
            ++n;
        }
    }
    return n;
}

/* chi_max of the renormalisation grid (integrals.c:855-880); 0 for outputs without one */
static double renormalisation_chi_max(const coffe_parameters_t *par, const coffe_background_t *bg,
                                      const coffe_spline_ops *ops)
{
    double zmax = -1.0;
    if (par->output_type == COFFE_REF_CORRFUNC || par->output_type == COFFE_REF_MULTIPOLES) {
        for (size_t i = 0; i < par->z_mean_len; ++i) {
            const double dz = (par->deltaz && i < par->deltaz_len) ? par->deltaz[i] : 0.0;
            zmax = fmax(zmax, par->z_mean[i] + dz);
        }
    } else if (par->output_type == COFFE_REF_AVERAGE_MULTIPOLES) {
        for (size_t i = 0; i < par->zmax_len; ++i) zmax = fmax(zmax, par->zmax[i]);
    }
    if (!(zmax >= 0.0) || !bg) return 0.0;
    return ops->interp_spline(&bg->comoving_distance, zmax);
}

int coffe_integrals_init(const coffe_parameters_t *par, const coffe_background_t *bg,
                         coffe_integral_array_t *integral)
{
    if (!par || !integral) return done(COFFE_EINVAL);
    coffe_spline_ops ops;
    int st = spline_ops(&ops);
    if (st != COFFE_OK) return done(st);
    st = coffe_integrals_free(integral);
    if (st != COFFE_OK) return done(st);
    if (par->pk_type != 0 && par->midpoint_approximation) return done(COFFE_EUNSUPPORTED);
    const int div = par->divergent != 0;
    /* the flat-sky lensing tables (integrals.c:915-962, 1082-1130, 1246-1300): the
       (1/2, -1/2) one, and for the multipoles the (1, l) and ((l-2k+3)/2, (l-2k+1)/2)
       ones of the even multipoles (the odd ones' terms are zero and their tables unused) */
    const int flat = par->flatsky_local_nonlocal || par->flatsky_nonlocal;
    const int flat_mp = par->output_type == COFFE_REF_MULTIPOLES;
    size_t extra = flat ? 1 : 0;
    if (flat && flat_mp && (par->multipole_values_len && !par->multipole_values)) return done(COFFE_EINVAL);
    for (size_t i = 0; flat && flat_mp && i < par->multipole_values_len; ++i) {
        const int l = par->multipole_values[i];
        if (l < 0) return done(COFFE_EINVAL);
        if (l % 2) continue;
        if (par->flatsky_nonlocal) {
            if (l > 4) return done(COFFE_EUNSUPPORTED);   /* the engine's (n, l) tables stop at l = 4 */
            ++extra;
        }
        if (par->flatsky_local_nonlocal) {
            if (l + 3 > 9) return done(COFFE_EUNSUPPORTED);
            extra += (size_t)(l / 2 + 1);
        }
    }
    nodes pk;
    if (nodes_of(&par->power_spectrum_norm, &pk) != COFFE_OK || par->bessel_bins <= 0) return done(COFFE_EINVAL);
    const size_t count = 8 + (size_t)div + extra;
    const size_t N = (size_t)par->bessel_bins, len = N + 43;
    coffe_integral_t *arr = (coffe_integral_t *)calloc(count, sizeof(coffe_integral_t));
    double *x = (double *)malloc(sizeof(double) * len), *y = (double *)malloc(sizeof(double) * len);
    if (!arr || !x || !y) {
        free(arr);
        free(x);
        free(y);
        return done(COFFE_ENOMEM);
    }
    integral->array = arr;
    integral->size = 0;
    for (int t = 0; t < 8 && st == COFFE_OK; ++t) {
        st = coffe_integral_table(STD_N[t], STD_L[t], N, pk.x, pk.y, pk.n, par->k_min_norm, par->k_max_norm, x, y);
        if (st != COFFE_OK) break;
        coffe_integral_t *I = &arr[integral->size++];
        I->n = STD_N[t];
        I->l = STD_L[t];
        if (ops.init_spline(&I->result, x, y, len, par->interp_method) != 0) st = COFFE_EINVAL;
    }
    free(x);
    free(y);
    if (st == COFFE_OK && div) {
        /* I_0^4 r^4 at the fixed separations, the r = 0 renormalisation with its first node
           replaced by (0, 0) (integrals.c:772-840), and the 201 x 201 grid of the
           renormalisation on [0, chi_max] (integrals.c:841-907) */
        const size_t ns = divergent_separations(NULL);
        const double chi_max = renormalisation_chi_max(par, bg, &ops);
        const size_t nchi = chi_max > 0.0 ? REF_REN_BINS + 1 : 0;
        double *sep = (double *)malloc(sizeof(double) * ns), *i40 = (double *)malloc(sizeof(double) * ns);
        double *r0 = (double *)malloc(sizeof(double) * ns);
        double *chi = (double *)malloc(sizeof(double) * (nchi + 1));
        double *ren = (double *)malloc(sizeof(double) * (nchi * nchi + 1));
        if (!sep || !i40 || !r0 || !chi || !ren) st = COFFE_ENOMEM;
        if (st == COFFE_OK) {
            divergent_separations(sep);
            for (size_t j = 0; j < nchi; ++j) chi[j] = (double)j / REF_REN_BINS * chi_max;
            st = coffe_divergent_tables(pk.x, pk.y, pk.n, par->k_min_norm, par->k_max_norm, sep, ns, chi, nchi, i40,
                                        r0, ren, NULL, NULL);
            /* as the reference (its GSL error handler is off), integrals stopped by the
               interval cap keep their values */
            if (st == COFFE_ENOCONV) st = COFFE_OK;
        }
        if (st == COFFE_OK) {
            coffe_integral_t *I = &arr[integral->size++];
            I->n = 4;
            I->l = 0;
            if (ops.init_spline(&I->result, sep, i40, ns, par->interp_method) != 0) st = COFFE_EINVAL;
            sep[0] = 0.0;
            r0[0] = 0.0;
            if (st == COFFE_OK &&
                ops.init_spline(&I->renormalization_zero_separation, sep, r0, ns, par->interp_method) != 0)
                st = COFFE_EINVAL;
            if (st == COFFE_OK && nchi &&
                ops.init_spline2d(&I->renormalization, chi, chi, ren, nchi, nchi,
                                  par->covariance_interpolation_method) != 0)
                st = COFFE_EINVAL;
        }
        free(sep);
        free(i40);
        free(r0);
        free(chi);
        free(ren);
    }
    if (st == COFFE_OK && flat) {
        double *fx = (double *)malloc(sizeof(double) * len), *fy = (double *)malloc(sizeof(double) * len);
        /* the tables (n, l, state) to add, each unless the array holds it already */
        int *tn = (int *)malloc(sizeof(int) * 3 * extra), *tl = tn ? tn + extra : NULL, *ts = tn ? tn + 2 * extra : NULL;
        size_t nt = 0;
        if (!fx || !fy || !tn) st = COFFE_ENOMEM;
        if (st == COFFE_OK) {
            tn[nt] = 1, tl[nt] = -1, ts[nt] = 1, ++nt;
            for (size_t i = 0; flat_mp && i < par->multipole_values_len; ++i) {
                const int l = par->multipole_values[i];
                if (l % 2) continue;
                if (par->flatsky_nonlocal) tn[nt] = 1, tl[nt] = l, ts[nt] = 0, ++nt;
                for (int k = 0; par->flatsky_local_nonlocal && k <= l / 2; ++k)
                    tn[nt] = l - 2 * k + 3, tl[nt] = l - 2 * k + 1, ts[nt] = 1, ++nt;
            }
        }
        for (size_t t = 0; t < nt && st == COFFE_OK; ++t) {
            if (coffe_find_integral(integral, tn[t], tl[t], ts[t], ts[t])) continue;
            st = coffe_integral_table_state(tn[t], tl[t], ts[t], ts[t], N, pk.x, pk.y, pk.n, par->k_min_norm,
                                            par->k_max_norm, fx, fy);
            if (st != COFFE_OK) break;
            /* the (1/2, -1/2) table is kept as sqrt(2/pi) 2 pi^2 r I(r) (integrals.c:956-962) */
            if (tn[t] == 1 && tl[t] == -1 && ts[t] == 1)
                for (size_t i = 0; i < len; ++i) fy[i] *= REF_FLAT_TSCALE;
            coffe_integral_t *I = &arr[integral->size++];
            I->n = tn[t];
            I->l = tl[t];
            I->state_n = I->state_l = ts[t];
            if (ops.init_spline(&I->result, fx, fy, len, par->interp_method) != 0) st = COFFE_EINVAL;
        }
        free(fx);
        free(fy);
        free(tn);
    }
    if (st != COFFE_OK) {
        coffe_integrals_free(integral);
        return done(st);
    }
    return done(COFFE_OK);
}

/* ---- the local terms (functions.c:89-1484) ------------------------------------------ */

typedef struct {
    coffe_mu_inputs in;
    unsigned terms, sources;
    int flatsky;
    int ntab, tab_n[9], tab_l[9];
    const coffe_integral_t *tab[9];
} local_ctx;

static int local_flags(const coffe_parameters_t *par)
{
    const coffe_correlation_contributions *c = &par->correlation_contrib;
    return c->den || c->rsd || c->d1 || c->d2 || c->g1 || c->g2 || c->g3;
}

static int local_setup(const coffe_parameters_t *par, const coffe_background_t *bg,
                       const coffe_integral_array_t *integral, local_ctx *lc)
{
    memset(lc, 0, sizeof(*lc));
    const coffe_correlation_contributions *c = &par->correlation_contrib;
    /* the derivative outputs and the CLASS spectra at z change the terms' form */
    if (par->b_derivative || par->f_derivative || par->b_tilde_derivative || par->f_tilde_derivative ||
        par->pk_type != 0)
        return COFFE_EUNSUPPORTED;
    unsigned src = (c->d2 ? COFFE_SOURCE_D2 : 0u) | (c->g1 ? COFFE_SOURCE_G1 : 0u) |
                   (c->g2 ? COFFE_SOURCE_G2 : 0u) | (c->g3 ? COFFE_SOURCE_G3 : 0u);
    unsigned t = coffe_terms_of(c->den != 0, c->rsd != 0, c->d1 != 0);
    if (src) {
        t |= COFFE_TERM_DG_DG;
        if (c->den) t |= COFFE_TERM_DEN_DG;
        if (c->rsd) t |= COFFE_TERM_RSD_DG;
        if (c->d1) t |= COFFE_TERM_D1_DG;
    }
    if (par->only_cross_correlations) {
        /* the reference skips the auto terms (functions.c:129-669), those of the d2 / g
           sources among them */
        t &= ~(unsigned)(COFFE_TERM_DEN_DEN | COFFE_TERM_RSD_RSD | COFFE_TERM_D1_D1);
        if (src) t |= COFFE_TERM_DG_MUTUAL;
    }
    lc->terms = t;
    lc->sources = src;
    lc->flatsky = par->flatsky_local != 0;
    if (!t) return COFFE_OK;
    if (!bg || !integral) return COFFE_EINVAL;
    /* the engine interpolates its tables with Akima, the reference's default method */
    if (par->interp_method != COFFE_REF_INTERP_AKIMA) return COFFE_EUNSUPPORTED;
    nodes chi, v;
    if (nodes_of(&bg->comoving_distance, &chi) != COFFE_OK) return COFFE_EINVAL;
    lc->in.nbg = (int)chi.n;
    lc->in.bg_z = chi.x;
    lc->in.bg_chi = chi.y;
    const coffe_interpolation *bgs[6] = {&bg->D1, &bg->f, &bg->conformal_Hz, &bg->G1, &bg->G2, &bg->a};
    const double **dst[6] = {&lc->in.bg_D1, &lc->in.bg_f, &lc->in.bg_cH, &lc->in.bg_G1, &lc->in.bg_G2,
                             &lc->in.bg_a};
    for (int i = 0; i < 6; ++i) {
        if (nodes_of(bgs[i], &v) != COFFE_OK) return COFFE_EINVAL;
        if (!same_nodes(&v, &chi)) return COFFE_EUNSUPPORTED;
        *dst[i] = v.y;
    }
    nodes b1, b2;
    if (nodes_of(&par->galaxy_bias1, &b1) != COFFE_OK || nodes_of(&par->galaxy_bias2, &b2) != COFFE_OK)
        return COFFE_EINVAL;
    if (!same_nodes(&b1, &b2)) return COFFE_EUNSUPPORTED;
    lc->in.nb = (int)b1.n;
    lc->in.bias_z = b1.x;
    lc->in.bias1 = b1.y;
    lc->in.bias2 = b2.y;
    for (int k = 0; k < 8; ++k) {
        const coffe_integral_t *I = coffe_find_integral(integral, STD_N[k], STD_L[k], 0, 0);
        if (!I || !I->result.spline) return COFFE_EINVAL;
        lc->tab[lc->ntab] = I;
        lc->tab_n[lc->ntab] = STD_N[k];
        lc->tab_l[lc->ntab] = STD_L[k];
        ++lc->ntab;
    }
    if (src) {
        nodes m[4];
        const coffe_interpolation *ms[4] = {&par->magnification_bias1, &par->magnification_bias2,
                                            &par->evolution_bias1, &par->evolution_bias2};
        for (int i = 0; i < 4; ++i) {
            if (nodes_of(ms[i], &m[i]) != COFFE_OK) return COFFE_EINVAL;
            if (!same_nodes(&m[i], &m[0])) return COFFE_EUNSUPPORTED;
        }
        lc->in.nmb = (int)m[0].n;
        lc->in.mb_z = m[0].x;
        lc->in.s1 = m[0].y;
        lc->in.s2 = m[1].y;
        lc->in.fevo1 = m[2].y;
        lc->in.fevo2 = m[3].y;
        lc->in.H0h = REF_H0 * par->h;
        lc->in.Omega_m = par->Omega0_cdm + par->Omega0_baryon;
        const coffe_integral_t *I = coffe_find_integral(integral, 4, 0, 0, 0);
        if (!I || !I->result.spline) return COFFE_EINVAL;
        const gsl_spline2d_view *r = (const gsl_spline2d_view *)I->renormalization.spline;
        if (!r || !r->xarr || !r->yarr || !r->zarr) return COFFE_EINVAL;
        /* the engine's grid: one chi node set for both axes, GSL's bicubic */
        if (r->obj.xsize != r->obj.ysize || memcmp(r->xarr, r->yarr, sizeof(double) * r->obj.xsize) != 0 ||
            par->covariance_interpolation_method != COFFE_REF_INTERP2D_BICUBIC)
            return COFFE_EUNSUPPORTED;
        lc->in.nren = (int)r->obj.xsize;
        lc->in.ren_chi = r->xarr;
        lc->in.ren = r->zarr;
        lc->tab[lc->ntab] = I;
        lc->tab_n[lc->ntab] = 4;
        lc->tab_l[lc->ntab] = 0;
        ++lc->ntab;
    }
    return COFFE_OK;
}

/*
 * Local multipoles (or, with a one-node rule, the correlation function) through the
 * engine's mu weights, contracted with the host's own interpolation of each table at
 * the separations.  Grid mode: out[z][l][r]; paired mode (nz == nsep): point i is
 * (zs[i], seps[i]) and out[i][l].  mu / wmu: an explicit rule, or NULL for the
 * REF_MU_NODES-node Gauss-Legendre one.  nl <= REF_MAX_L.
 */
static int local_eval(const local_ctx *lc, const coffe_spline_ops *ops, int paired, int nz, const double *zs,
                      int nl, const int *ls, int nsep, const double *seps, int nmu, const double *mu,
                      const double *wmu, double *out)
{
    const size_t npt = paired ? (size_t)nz : (size_t)nz * (size_t)nsep;
    if (!lc->terms) {
        memset(out, 0, sizeof(double) * npt * (size_t)nl);
        return COFFE_OK;
    }
    coffe_engine *e = coffe_engine_default();
    if (!e) return COFFE_ENODEVICE;
    const int ntab = lc->ntab, ren = (lc->terms & COFFE_TERM_DG_DG) != 0;
    const size_t per = paired ? 1 : (size_t)nsep;
    const size_t nA = (size_t)nz * (size_t)nl * (size_t)ntab * per, nR = (size_t)nz * (size_t)nl * per;
    double *tv = (double *)malloc(sizeof(double) * (size_t)ntab * (size_t)nsep);
    double *hA = (double *)malloc(sizeof(double) * nA), *hR = (double *)malloc(sizeof(double) * nR);
    void *dA = NULL, *dR = NULL;
    int st = (tv && hA && hR) ? COFFE_OK : COFFE_ENOMEM;
    if (st == COFFE_OK) {
        for (int t = 0; t < ntab; ++t)
            for (int r = 0; r < nsep; ++r) tv[(size_t)t * nsep + r] = ops->interp_spline(&lc->tab[t]->result, seps[r]);
        hipSetDevice(e->device);
        st = ce_dev_alloc(&dA, sizeof(double) * nA);
        if (st == COFFE_OK && ren) st = ce_dev_alloc(&dR, sizeof(double) * nR);
    }
    if (st == COFFE_OK) {
        st = mu ? coffe_mu_weights_rule(e, &lc->in, lc->terms, lc->sources, lc->flatsky, paired, nz, zs, nl, ls, ntab,
                                        lc->tab_n, lc->tab_l, nsep, seps, nmu, mu, wmu, (double *)dA, (double *)dR,
                                        NULL)
                : coffe_mu_weights_ex(e, &lc->in, lc->terms, lc->sources, lc->flatsky, paired, nz, zs, nl, ls, ntab,
                                      lc->tab_n, lc->tab_l, nsep, seps, REF_MU_NODES, (double *)dA, (double *)dR,
                                      NULL);
    }
    if (st == COFFE_OK) st = coffe_engine_synchronize(e);
    if (st == COFFE_OK) st = ce_hip_ok(hipMemcpy(hA, dA, sizeof(double) * nA, hipMemcpyDeviceToHost));
    if (st == COFFE_OK && ren) st = ce_hip_ok(hipMemcpy(hR, dR, sizeof(double) * nR, hipMemcpyDeviceToHost));
    if (st == COFFE_OK) {
        for (int z = 0; z < nz; ++z)
            for (int l = 0; l < nl; ++l)
                for (size_t r = 0; r < per; ++r) {
                    const int sr = paired ? z : (int)r;   /* separation index of the point */
                    const size_t zl = (size_t)z * nl + l;
                    double v = ren ? hR[zl * per + r] : 0.0;
                    for (int t = 0; t < ntab; ++t) v += hA[(zl * ntab + t) * per + r] * tv[(size_t)t * nsep + sr];
                    out[zl * per + r] = v;
                }
    }
    if (dA) hipFree(dA);
    if (dR) hipFree(dR);
    free(tv);
    free(hA);
    free(hR);
    return st;
}

/* local_eval in groups of at most REF_MAX_L multipoles: out[z][l][r] (grid) or out[i][l] */
static int local_eval_ls(const local_ctx *lc, const coffe_spline_ops *ops, int paired, int nz, const double *zs,
                         int nl, const int *ls, int nsep, const double *seps, double *out)
{
    const size_t per = paired ? 1 : (size_t)nsep;
    double *part = (double *)malloc(sizeof(double) * (size_t)nz * per * REF_MAX_L);
    if (!part) return COFFE_ENOMEM;
    int st = COFFE_OK;
    for (int l0 = 0; l0 < nl && st == COFFE_OK; l0 += REF_MAX_L) {
        const int m = nl - l0 < REF_MAX_L ? nl - l0 : REF_MAX_L;
        st = local_eval(lc, ops, paired, nz, zs, m, ls + l0, nsep, seps, 0, NULL, NULL, part);
        if (st != COFFE_OK) break;
        for (int z = 0; z < nz; ++z)
            for (int j = 0; j < m; ++j)
                memcpy(out + ((size_t)z * nl + l0 + j) * per, part + ((size_t)z * m + j) * per, sizeof(double) * per);
    }
    free(part);
    return st;
}

/* n-point Gauss-Legendre rule on [0, 1] (Newton on the three-term recurrence) */
static void gauss_legendre01(int n, double *x, double *w)
{
    for (int i = 0; i < n; ++i) {
        double t = cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = t;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (t * p1 - p0) / (t * t - 1.0);
            const double dt = p1 / dp;
            t -= dt;
            if (fabs(dt) < 1e-16) break;
        }
        x[i] = 0.5 * (1.0 - t);
        w[i] = 1.0 / ((1.0 - t * t) * dp * dp);
    }
}

/*
 * Multipoles averaged over the redshift bin [zmin, zmax] (signal.c:249-311, 543-550):
 * for separation r, z runs over [z1, z2], z1 = z(chi(zmin) + r/2), z2 = z(chi(zmax) -
 * r/2), and the average is int_0^1 dv M_l(z(v), r) h H0 / H(z(v)), z(v) = z1 + (z2 - z1)
 * v, on a REF_Z_NODES-node Gauss-Legendre rule in v with the multipoles at exactly the
 * (z(v), r) pairs (the engine's paired mode).  out[l][r].
 */
static int average_eval(const local_ctx *lc, const coffe_spline_ops *ops, const coffe_parameters_t *par,
                        const coffe_background_t *bg, double zmin, double zmax, int nsep, const double *seps, int nl,
                        const int *ls, double *out)
{
    if (!lc->terms) {
        memset(out, 0, sizeof(double) * (size_t)nl * nsep);
        return COFFE_OK;
    }
    const int nv = REF_Z_NODES;
    const size_t np = (size_t)nsep * nv;
    double v[REF_Z_NODES], w[REF_Z_NODES];
    gauss_legendre01(nv, v, w);
    double *zn = (double *)malloc(sizeof(double) * np), *sp = (double *)malloc(sizeof(double) * np);
    double *wz = (double *)malloc(sizeof(double) * np), *m = (double *)malloc(sizeof(double) * np * nl);
    int st = (zn && sp && wz && m) ? COFFE_OK : COFFE_ENOMEM;
    const double hH0 = par->h * REF_H0;
    const double chi_lo = st == COFFE_OK ? ops->interp_spline(&bg->comoving_distance, zmin) : 0.0;
    const double chi_hi = st == COFFE_OK ? ops->interp_spline(&bg->comoving_distance, zmax) : 0.0;
    for (int r = 0; r < nsep && st == COFFE_OK; ++r) {
        const double z1 = ops->interp_spline(&bg->z_as_chi, chi_lo + seps[r] / 2);
        const double z2 = ops->interp_spline(&bg->z_as_chi, chi_hi - seps[r] / 2);
        if (!(z2 > z1)) {
            st = COFFE_EDOMAIN;   /* the separation exceeds the comoving depth of the bin */
            break;
        }
        for (int j = 0; j < nv; ++j) {
            const size_t i = (size_t)r * nv + j;
            zn[i] = z1 + (z2 - z1) * v[j];
            sp[i] = seps[r];
            wz[i] = w[j] * hH0 / ops->interp_spline(&bg->Hz, zn[i]);
        }
    }
    if (st == COFFE_OK) st = local_eval_ls(lc, ops, 1, (int)np, zn, nl, ls, (int)np, sp, m);
    if (st == COFFE_OK)
        for (int l = 0; l < nl; ++l)
            for (int r = 0; r < nsep; ++r) {
                double acc = 0.0;
                for (int j = 0; j < nv; ++j) {
                    const size_t i = (size_t)r * nv + j;
                    acc += wz[i] * m[i * nl + l];
                }
                out[(size_t)l * nsep + r] = acc;
            }
    free(zn);
    free(sp);
    free(wz);
    free(m);
    return st;
}

/* ---- the flat-sky lensing terms (signal.c:556-771) ------------------------------------ */

/* the terms' values from the integral array's own tables, on the GPU
   (coffe_flat_lensing_multipoles / coffe_flat_lensing_corrfunc): lensing-lensing
   (want_ll) and density-lensing (want_dl), as signal.c adds them (the multipoles times
   2 pi^2); corrfunc: at mus[nl] in place of the multipoles ls[nl] */
static int flat_eval(const coffe_parameters_t *par, const coffe_background_t *bg,
                     const coffe_integral_array_t *integral, int corrfunc, int want_ll, int want_dl, int nz,
                     const double *zs, int nl, const int *ls, const double *mus, int nsep, const double *seps,
                     double *out)
{
    if (!bg || !integral) return COFFE_EINVAL;
    if (par->interp_method != COFFE_REF_INTERP_AKIMA || par->pk_type != 0) return COFFE_EUNSUPPORTED;
    coffe_spline_ops ops;
    int st = spline_ops(&ops);
    if (st != COFFE_OK) return st;
    nodes bgc, bgd;
    if (nodes_of(&bg->comoving_distance, &bgc) != COFFE_OK || nodes_of(&bg->D1, &bgd) != COFFE_OK) return COFFE_EINVAL;
    if (!same_nodes(&bgc, &bgd)) return COFFE_EUNSUPPORTED;
    double *zpar = (double *)malloc(sizeof(double) * 7 * (size_t)nz);
    if (!zpar) return COFFE_ENOMEM;
    for (int i = 0; i < nz; ++i) {
        double *zp = zpar + 7 * i;
        zp[0] = zs[i];
        zp[1] = ops.interp_spline(&bg->comoving_distance, zs[i]);
        zp[2] = ops.interp_spline(&bg->D1, zs[i]);
        zp[3] = ops.interp_spline(&par->magnification_bias1, zs[i]);
        zp[4] = ops.interp_spline(&par->magnification_bias2, zs[i]);
        zp[5] = ops.interp_spline(&par->galaxy_bias1, zs[i]);
        zp[6] = ops.interp_spline(&par->galaxy_bias2, zs[i]);
    }
    const double Om = par->Omega0_cdm + par->Omega0_baryon, H0h = REF_H0 * par->h;
    if (corrfunc) {
        /* the (1/2, -1/2) table, kept scaled by sqrt(2/pi) 2 pi^2 in the array */
        const coffe_integral_t *I = coffe_find_integral(integral, 1, -1, 1, 1);
        nodes t;
        if (!I || nodes_of(&I->result, &t) != COFFE_OK) {
            free(zpar);
            return COFFE_EINVAL;
        }
        double *ty = (double *)malloc(sizeof(double) * t.n);
        if (!ty) st = COFFE_ENOMEM;
        for (size_t i = 0; st == COFFE_OK && i < t.n; ++i) ty[i] = t.y[i] / REF_FLAT_TSCALE;
        if (st == COFFE_OK)
            st = coffe_flat_lensing_corrfunc(nz, zpar, nl, mus, nsep, seps, (int)bgc.n, bgc.x, bgc.y, bgd.y, (int)t.n,
                                             t.x, ty, Om, H0h, want_ll, want_dl, 128, out);
        free(ty);
        free(zpar);
        return st;
    }
    /* the multipoles: the tables of every even multipole, concatenated */
    int64_t *off = (int64_t *)calloc((size_t)nl * (1 + CE_FL_MAXK), sizeof(int64_t));
    int *lens = (int *)calloc((size_t)nl * (1 + CE_FL_MAXK), sizeof(int));
    const nodes **tabs = (const nodes **)calloc((size_t)nl * (1 + CE_FL_MAXK), sizeof(nodes *));
    nodes *store = (nodes *)calloc((size_t)nl * (1 + CE_FL_MAXK), sizeof(nodes));
    if (!off || !lens || !tabs || !store) st = COFFE_ENOMEM;
    int64_t total = 0;
    for (int il = 0; il < nl && st == COFFE_OK; ++il) {
        const int l = ls[il];
        if (l % 2) continue;
        for (int slot = 0; slot < 1 + CE_FL_MAXK && st == COFFE_OK; ++slot) {
            const int k = slot - 1;
            const coffe_integral_t *I = NULL;
            if (slot == 0 && want_ll) I = coffe_find_integral(integral, 1, l, 0, 0);
            else if (slot > 0 && want_dl && k <= l / 2) I = coffe_find_integral(integral, l - 2 * k + 3, l - 2 * k + 1, 1, 1);
            else continue;
            const size_t q = (size_t)il * (1 + CE_FL_MAXK) + (size_t)slot;
            if (!I || nodes_of(&I->result, &store[q]) != COFFE_OK) {
                st = COFFE_EINVAL;
                break;
            }
            tabs[q] = &store[q];
            off[q] = total;
            lens[q] = (int)store[q].n;
            total += (int64_t)store[q].n;
        }
    }
    double *tx = NULL, *ty = NULL;
    int64_t *ll_off = NULL, *dl_off = NULL;
    int *ll_len = NULL, *dl_len = NULL;
    if (st == COFFE_OK) {
        tx = (double *)malloc(sizeof(double) * (size_t)(total > 0 ? total : 1));
        ty = (double *)malloc(sizeof(double) * (size_t)(total > 0 ? total : 1));
        ll_off = (int64_t *)calloc((size_t)nl, sizeof(int64_t));
        ll_len = (int *)calloc((size_t)nl, sizeof(int));
        dl_off = (int64_t *)calloc((size_t)nl * CE_FL_MAXK, sizeof(int64_t));
        dl_len = (int *)calloc((size_t)nl * CE_FL_MAXK, sizeof(int));
        if (!tx || !ty || !ll_off || !ll_len || !dl_off || !dl_len) st = COFFE_ENOMEM;
    }
    for (int il = 0; il < nl && st == COFFE_OK; ++il)
        for (int slot = 0; slot < 1 + CE_FL_MAXK; ++slot) {
            const size_t q = (size_t)il * (1 + CE_FL_MAXK) + (size_t)slot;
            if (!tabs[q]) continue;
            memcpy(tx + off[q], tabs[q]->x, sizeof(double) * tabs[q]->n);
            memcpy(ty + off[q], tabs[q]->y, sizeof(double) * tabs[q]->n);
            if (slot == 0) {
                ll_off[il] = off[q];
                ll_len[il] = lens[q];
            } else {
                dl_off[(size_t)il * CE_FL_MAXK + slot - 1] = off[q];
                dl_len[(size_t)il * CE_FL_MAXK + slot - 1] = lens[q];
            }
        }
    if (st == COFFE_OK)
        st = coffe_flat_lensing_multipoles(nz, zpar, nl, ls, nsep, seps, (int)bgc.n, bgc.x, bgc.y, bgd.y, total, tx, ty,
                                           ll_off, ll_len, dl_off, dl_len, Om, H0h, want_ll, want_dl, 128, out);
    free(off);
    free(lens);
    free(tabs);
    free(store);
    free(tx);
    free(ty);
    free(ll_off);
    free(ll_len);
    free(dl_off);
    free(dl_len);
    free(zpar);
    return st;
}

/* ---- the full-sky integrated terms (signal.c:556-771) -------------------------------- */

#define REF_INTG_MU_NODES 64   /* theta nodes (mu = cos theta) of the integrated multipoles */

/* the contributions of the integrated kernel and whether any full-sky term remains
   (lensing-lensing and density/RSD/d1-lensing left to the flat-sky forms under the
   flat-sky switches, signal.c:612-628, 704-736) */
static unsigned intg_flags(const coffe_parameters_t *par, int *any)
{
    const coffe_correlation_contributions *c = &par->correlation_contrib;
    unsigned f = (c->den ? COFFE_INTEGRATED_DEN : 0u) | (c->rsd ? COFFE_INTEGRATED_RSD : 0u) |
                 (c->d1 ? COFFE_INTEGRATED_D1 : 0u) | (c->d2 ? COFFE_INTEGRATED_D2 : 0u) |
                 (c->g1 ? COFFE_INTEGRATED_G1 : 0u) | (c->g2 ? COFFE_INTEGRATED_G2 : 0u) |
                 (c->g3 ? COFFE_INTEGRATED_G3 : 0u) | (c->len ? COFFE_INTEGRATED_LEN : 0u) |
                 (c->g4 ? COFFE_INTEGRATED_G4 : 0u) | (c->g5 ? COFFE_INTEGRATED_G5 : 0u);
    const int oc = par->only_cross_correlations != 0, fl = par->flatsky_nonlocal != 0,
              fdl = par->flatsky_local_nonlocal != 0;
    if (oc) f |= COFFE_INTEGRATED_ONLY_CROSS;
    if (fl) f |= COFFE_INTEGRATED_FLAT_LL;
    if (fdl) f |= COFFE_INTEGRATED_FLAT_DL;
    const int local = local_flags(par), g45 = c->g4 || c->g5;
    const int dl_local = (c->den || c->rsd || c->d1) && !fdl;
    const int dbl = (c->len && !oc && !fl) || (g45 && !oc) || (c->len && g45) || (c->g4 && c->g5);
    const int sgl = (c->len && (dl_local || c->d2 || c->g1 || c->g2 || c->g3)) || (g45 && local);
    *any = dbl || sgl;
    return f;
}

/* the single- and double-integrated correlation function at every (z, mu, sep) from the
   integral array's own tables (coffe_integrated_corrfunc): out[z][mu][r][2] */
static int intg_corrfunc(const coffe_parameters_t *par, const coffe_background_t *bg,
                         const coffe_integral_array_t *integral, unsigned flags, int nz, const double *zs, int nmu,
                         const double *mus, int nsep, const double *seps, double *out)
{
    if (!bg || !integral) return COFFE_EINVAL;
    if (par->interp_method != COFFE_REF_INTERP_AKIMA || par->pk_type != 0 || par->b_derivative ||
        par->f_derivative || par->b_tilde_derivative || par->f_tilde_derivative)
        return COFFE_EUNSUPPORTED;
    coffe_mu_inputs in;
    memset(&in, 0, sizeof(in));
    nodes chi, v;
    if (nodes_of(&bg->comoving_distance, &chi) != COFFE_OK) return COFFE_EINVAL;
    in.nbg = (int)chi.n;
    in.bg_z = chi.x;
    in.bg_chi = chi.y;
    const coffe_interpolation *bgs[6] = {&bg->D1, &bg->f, &bg->conformal_Hz, &bg->G1, &bg->G2, &bg->a};
    const double **dst[6] = {&in.bg_D1, &in.bg_f, &in.bg_cH, &in.bg_G1, &in.bg_G2, &in.bg_a};
    for (int i = 0; i < 6; ++i) {
        if (nodes_of(bgs[i], &v) != COFFE_OK) return COFFE_EINVAL;
        if (!same_nodes(&v, &chi)) return COFFE_EUNSUPPORTED;
        *dst[i] = v.y;
    }
    nodes b1, b2, m[4];
    if (nodes_of(&par->galaxy_bias1, &b1) != COFFE_OK || nodes_of(&par->galaxy_bias2, &b2) != COFFE_OK)
        return COFFE_EINVAL;
    if (!same_nodes(&b1, &b2)) return COFFE_EUNSUPPORTED;
    in.nb = (int)b1.n;
    in.bias_z = b1.x;
    in.bias1 = b1.y;
    in.bias2 = b2.y;
    const coffe_interpolation *ms[4] = {&par->magnification_bias1, &par->magnification_bias2, &par->evolution_bias1,
                                        &par->evolution_bias2};
    for (int i = 0; i < 4; ++i) {
        if (nodes_of(ms[i], &m[i]) != COFFE_OK) return COFFE_EINVAL;
        if (!same_nodes(&m[i], &m[0])) return COFFE_EUNSUPPORTED;
    }
    in.nmb = (int)m[0].n;
    in.mb_z = m[0].x;
    in.s1 = m[0].y;
    in.s2 = m[1].y;
    in.fevo1 = m[2].y;
    in.fevo2 = m[3].y;
    in.H0h = REF_H0 * par->h;
    in.Omega_m = par->Omega0_cdm + par->Omega0_baryon;
    /* the 8 tables and, with g4 / g5, the divergent one and its renormalisations */
    int tn[9], tl[9], tlen[9], nt = 0;
    nodes tabs[9];
    for (int k = 0; k < 8; ++k) {
        const coffe_integral_t *I = coffe_find_integral(integral, STD_N[k], STD_L[k], 0, 0);
        if (!I || nodes_of(&I->result, &tabs[nt]) != COFFE_OK) return COFFE_EINVAL;
        tn[nt] = STD_N[k];
        tl[nt] = STD_L[k];
        tlen[nt] = (int)tabs[nt].n;
        ++nt;
    }
    nodes r0 = {NULL, NULL, 0};
    if (flags & (COFFE_INTEGRATED_G4 | COFFE_INTEGRATED_G5)) {
        const coffe_integral_t *I = coffe_find_integral(integral, 4, 0, 0, 0);
        if (!I || nodes_of(&I->result, &tabs[nt]) != COFFE_OK ||
            nodes_of(&I->renormalization_zero_separation, &r0) != COFFE_OK)
            return COFFE_EINVAL;
        const gsl_spline2d_view *r = (const gsl_spline2d_view *)I->renormalization.spline;
        if (!r || !r->xarr || !r->yarr || !r->zarr) return COFFE_EINVAL;
        if (r->obj.xsize != r->obj.ysize || memcmp(r->xarr, r->yarr, sizeof(double) * r->obj.xsize) != 0 ||
            par->covariance_interpolation_method != COFFE_REF_INTERP2D_BICUBIC)
            return COFFE_EUNSUPPORTED;
        in.nren = (int)r->obj.xsize;
        in.ren_chi = r->xarr;
        in.ren = r->zarr;
        tn[nt] = 4;
        tl[nt] = 0;
        tlen[nt] = (int)tabs[nt].n;
        ++nt;
    }
    size_t total = 0;
    for (int t = 0; t < nt; ++t) total += tabs[t].n;
    double *tx = (double *)malloc(sizeof(double) * total), *ty = (double *)malloc(sizeof(double) * total);
    int st = (tx && ty) ? COFFE_OK : COFFE_ENOMEM;
    for (int t = 0, o = 0; st == COFFE_OK && t < nt; o += tlen[t], ++t) {
        memcpy(tx + o, tabs[t].x, sizeof(double) * tabs[t].n);
        memcpy(ty + o, tabs[t].y, sizeof(double) * tabs[t].n);
    }
    if (st == COFFE_OK)
        st = coffe_integrated_corrfunc(&in, flags, nt, tn, tl, tlen, tx, ty, (int)r0.n, r0.x, r0.y, nz, zs, nmu, mus,
                                       nsep, seps, NULL, out);
    free(tx);
    free(ty);
    return st;
}

/* one component (0 single, 1 double, -1 both) of the integrated multipoles: out[z][l][r]
   = (2 l + 1) / 2 int dmu P_l(mu) xi(mu), theta (mu = cos theta) on a REF_INTG_MU_NODES
   Gauss-Legendre rule */
static int intg_multipoles(const coffe_parameters_t *par, const coffe_background_t *bg,
                           const coffe_integral_array_t *integral, unsigned flags, int which, int nz, const double *zs,
                           int nl, const int *ls, int nsep, const double *seps, double *out)
{
    const int nm = REF_INTG_MU_NODES;
    double t[REF_INTG_MU_NODES], w[REF_INTG_MU_NODES], mu[REF_INTG_MU_NODES], wm[REF_INTG_MU_NODES];
    gauss_legendre01(nm, t, w);
    for (int j = 0; j < nm; ++j) {
        const double th = M_PI * t[j];
        mu[j] = cos(th);
        wm[j] = M_PI * w[j] * sin(th);
    }
    double *xi = (double *)malloc(sizeof(double) * 2 * (size_t)nz * nm * nsep);
    if (!xi) return COFFE_ENOMEM;
    int st = intg_corrfunc(par, bg, integral, flags, nz, zs, nm, mu, nsep, seps, xi);
    for (int z = 0; st == COFFE_OK && z < nz; ++z)
        for (int il = 0; il < nl; ++il) {
            const int l = ls[il];
            for (int r = 0; r < nsep; ++r) {
                double acc = 0.0;
                for (int j = 0; j < nm; ++j) {
                    /* P_l(mu) by the three-term recurrence */
                    double p0 = 1.0, p1 = mu[j];
                    double pl = l == 0 ? 1.0 : mu[j];
                    for (int k = 2; k <= l; ++k) {
                        pl = ((2.0 * k - 1.0) * mu[j] * p1 - (k - 1.0) * p0) / k;
                        p0 = p1;
                        p1 = pl;
                    }
                    const double *q = xi + 2 * (((size_t)z * nm + j) * nsep + r);
                    const double val = which < 0 ? q[0] + q[1] : q[which];
                    acc += wm[j] * pl * val;
                }
                out[((size_t)z * nl + il) * nsep + r] = (2.0 * l + 1.0) / 2.0 * acc;
            }
        }
    free(xi);
    return st;
}

/* the integrated terms of the multipoles averaged over [zmin, zmax] (signal.c:312-470,
   674-682, 759-767): their correlation function (the flat-sky forms flat_ll / flat_dl
   and the full-sky kernel's component `which`: 0 single, 1 double, -1 both) at the
   redshift nodes of average_eval and the theta nodes of intg_multipoles, projected on
   P_l and averaged with average_eval's weights; out[l][r] */
static int intg_average(const coffe_parameters_t *par, const coffe_background_t *bg,
                        const coffe_integral_array_t *integral, const coffe_spline_ops *ops, int which, int flat_ll,
                        int flat_dl, double zmin, double zmax, int nsep, const double *seps, int nl, const int *ls,
                        double *out)
{
    const int nv = REF_Z_NODES, nm = REF_INTG_MU_NODES;
    double v[REF_Z_NODES], w[REF_Z_NODES], t[REF_INTG_MU_NODES], wt[REF_INTG_MU_NODES];
    double mu[REF_INTG_MU_NODES], wm[REF_INTG_MU_NODES], zn[REF_Z_NODES], wz[REF_Z_NODES];
    gauss_legendre01(nv, v, w);
    gauss_legendre01(nm, t, wt);
    for (int j = 0; j < nm; ++j) {
        mu[j] = cos(M_PI * t[j]);
        wm[j] = M_PI * wt[j] * sin(M_PI * t[j]);
    }
    int any = 0;
    const unsigned f = intg_flags(par, &any);
    double *xi = (double *)malloc(sizeof(double) * 2 * (size_t)nv * nm);
    double *fl = (double *)malloc(sizeof(double) * (size_t)nv * nm);
    double *tot = (double *)malloc(sizeof(double) * (size_t)nv * nm);
    int st = (xi && fl && tot) ? COFFE_OK : COFFE_ENOMEM;
    const double hH0 = par->h * REF_H0;
    const double chi_lo = ops->interp_spline(&bg->comoving_distance, zmin);
    const double chi_hi = ops->interp_spline(&bg->comoving_distance, zmax);
    for (int r = 0; r < nsep && st == COFFE_OK; ++r) {
        const double z1 = ops->interp_spline(&bg->z_as_chi, chi_lo + seps[r] / 2);
        const double z2 = ops->interp_spline(&bg->z_as_chi, chi_hi - seps[r] / 2);
        if (!(z2 > z1)) {
            st = COFFE_EDOMAIN;
            break;
        }
        for (int j = 0; j < nv; ++j) {
            zn[j] = z1 + (z2 - z1) * v[j];
            wz[j] = w[j] * hH0 / ops->interp_spline(&bg->Hz, zn[j]);
        }
        memset(tot, 0, sizeof(double) * (size_t)nv * nm);
        if (flat_ll || flat_dl) {
            st = flat_eval(par, bg, integral, 1, flat_ll, flat_dl, nv, zn, nm, NULL, mu, 1, &seps[r], fl);
            for (size_t q = 0; st == COFFE_OK && q < (size_t)nv * nm; ++q) tot[q] += fl[q];
        }
        if (st == COFFE_OK && any) {
            st = intg_corrfunc(par, bg, integral, f, nv, zn, nm, mu, 1, &seps[r], xi);
            for (size_t q = 0; st == COFFE_OK && q < (size_t)nv * nm; ++q)
                tot[q] += which < 0 ? xi[2 * q] + xi[2 * q + 1] : xi[2 * q + which];
        }
        for (int il = 0; st == COFFE_OK && il < nl; ++il) {
            const int l = ls[il];
            double acc = 0.0;
            for (int m = 0; m < nm; ++m) {
                double p0 = 1.0, p1 = mu[m], pl = l == 0 ? 1.0 : mu[m];
                for (int k = 2; k <= l; ++k) {
                    pl = ((2.0 * k - 1.0) * mu[m] * p1 - (k - 1.0) * p0) / k;
                    p0 = p1;
                    p1 = pl;
                }
                double s_ = 0.0;
                for (int j = 0; j < nv; ++j) s_ += wz[j] * tot[(size_t)j * nm + m];
                acc += wm[m] * pl * s_;
            }
            out[(size_t)il * nsep + r] = (2.0 * l + 1.0) / 2.0 * acc;
        }
    }
    free(xi);
    free(fl);
    free(tot);
    return st;
}

/* ---- signal.c ------------------------------------------------------------------------ */

double coffe_integrate(coffe_parameters_t *par, coffe_background_t *bg, coffe_integral_array_t *integral,
                       double z_mean, double z_min, double z_max, double sep, double mu, int l, int flag_integral,
                       int flag_output)
{
    if (!par) {
        done(COFFE_EINVAL);
        return NAN;
    }
    const coffe_correlation_contributions *c = &par->correlation_contrib;
    const int local = local_flags(par), nonlocal = c->len || c->g4 || c->g5;
    int st = COFFE_OK;
    double result = 0.0;
    switch (flag_integral) {
    case COFFE_REF_NONINTEGRATED: {
        if (!local) break;
        if (flag_output != COFFE_REF_CORRFUNC && flag_output != COFFE_REF_MULTIPOLES &&
            flag_output != COFFE_REF_AVERAGE_MULTIPOLES)
            break;   /* signal.c:552: nothing for the other outputs */
        coffe_spline_ops ops;
        local_ctx lc;
        st = spline_ops(&ops);
        if (st == COFFE_OK) st = local_setup(par, bg, integral, &lc);
        if (st != COFFE_OK) break;
        if (flag_output == COFFE_REF_CORRFUNC) {
            /* the correlation function at mu: one node of weight 2 and l = 0 */
            const double w2 = 2.0;
            const int l0 = 0;
            st = local_eval(&lc, &ops, 0, 1, &z_mean, 1, &l0, 1, &sep, 1, &mu, &w2, &result);
        } else if (flag_output == COFFE_REF_MULTIPOLES) {
            st = local_eval(&lc, &ops, 0, 1, &z_mean, 1, &l, 1, &sep, 0, NULL, NULL, &result);
        } else {
            st = average_eval(&lc, &ops, par, bg, z_min, z_max, 1, &sep, 1, &l, &result);
        }
        break;
    }
    case COFFE_REF_SINGLE_INTEGRATED: {
        /* signal.c:556-683: lensing, g4 or g5 with a local term.  Density-lensing takes its
           flat-sky form with flatsky_local_nonlocal (RSD- and Doppler-lensing vanish
           there); the rest is the full-sky single integral */
        if (!(nonlocal && local)) break;
        if (flag_output == COFFE_REF_AVERAGE_MULTIPOLES) {
            coffe_spline_ops ops;
            st = spline_ops(&ops);
            if (st == COFFE_OK && !bg) st = COFFE_EINVAL;
            if (st == COFFE_OK)
                st = intg_average(par, bg, integral, &ops, 0, 0, c->len && c->den && par->flatsky_local_nonlocal,
                                  z_min, z_max, 1, &sep, 1, &l, &result);
            break;
        }
        if (flag_output != COFFE_REF_CORRFUNC && flag_output != COFFE_REF_MULTIPOLES) break;
        const int cf = flag_output == COFFE_REF_CORRFUNC;
        if (c->len && c->den && par->flatsky_local_nonlocal)
            st = cf ? flat_eval(par, bg, integral, 1, 0, 1, 1, &z_mean, 1, NULL, &mu, 1, &sep, &result)
                    : flat_eval(par, bg, integral, 0, 0, 1, 1, &z_mean, 1, &l, NULL, 1, &sep, &result);
        int any = 0;
        const unsigned f = intg_flags(par, &any);
        double part = 0.0, both[2];
        if (st == COFFE_OK && any) {
            if (cf) {
                st = intg_corrfunc(par, bg, integral, f, 1, &z_mean, 1, &mu, 1, &sep, both);
                part = both[0];
            } else {
                st = intg_multipoles(par, bg, integral, f, 0, 1, &z_mean, 1, &l, 1, &sep, &part);
            }
        }
        result += part;
        break;
    }
    case COFFE_REF_DOUBLE_INTEGRATED: {
        /* signal.c:685-771: lensing, g4 and g5 with each other; lensing-lensing takes its
           flat-sky form with flatsky_nonlocal (not with only the cross terms: an auto term) */
        if (!nonlocal) break;
        if (flag_output == COFFE_REF_AVERAGE_MULTIPOLES) {
            coffe_spline_ops ops;
            st = spline_ops(&ops);
            if (st == COFFE_OK && !bg) st = COFFE_EINVAL;
            if (st == COFFE_OK)
                st = intg_average(par, bg, integral, &ops, 1,
                                  c->len && par->flatsky_nonlocal && !par->only_cross_correlations, 0, z_min, z_max, 1,
                                  &sep, 1, &l, &result);
            break;
        }
        if (flag_output != COFFE_REF_CORRFUNC && flag_output != COFFE_REF_MULTIPOLES) break;
        const int cf = flag_output == COFFE_REF_CORRFUNC;
        if (c->len && par->flatsky_nonlocal && !par->only_cross_correlations)
            st = cf ? flat_eval(par, bg, integral, 1, 1, 0, 1, &z_mean, 1, NULL, &mu, 1, &sep, &result)
                    : flat_eval(par, bg, integral, 0, 1, 0, 1, &z_mean, 1, &l, NULL, 1, &sep, &result);
        int any = 0;
        const unsigned f = intg_flags(par, &any);
        double part = 0.0, both[2];
        if (st == COFFE_OK && any) {
            if (cf) {
                st = intg_corrfunc(par, bg, integral, f, 1, &z_mean, 1, &mu, 1, &sep, both);
                part = both[1];
            } else {
                st = intg_multipoles(par, bg, integral, f, 1, 1, &z_mean, 1, &l, 1, &sep, &part);
            }
        }
        result += part;
        break;
    }
    default:
        st = COFFE_EINVAL;
    }
    done(st);
    return st == COFFE_OK ? result : NAN;
}

/* ---- multipoles.c / average_multipoles.c -------------------------------------------- */

int coffe_multipoles_free(coffe_multipoles_array_t *mp)
{
    if (!mp) return done(COFFE_EINVAL);
    free(mp->array);
    mp->array = NULL;
    mp->size = 0;
    return done(COFFE_OK);
}

int coffe_average_multipoles_free(coffe_average_multipoles_array_t *ramp)
{
    if (!ramp) return done(COFFE_EINVAL);
    free(ramp->array);
    ramp->array = NULL;
    ramp->size = 0;
    return done(COFFE_OK);
}

/* the checks shared by the two multipole outputs; *ll / *dl (may be NULL: refused)
   receive the flat-sky lensing terms to add (signal.c:612-628, 704-736) */
static int multipole_setup(const coffe_parameters_t *par, const coffe_background_t *bg,
                           const coffe_integral_array_t *integral, coffe_spline_ops *ops, local_ctx *lc, int *ll,
                           int *dl)
{
    const coffe_correlation_contributions *c = &par->correlation_contrib;
    if (ll) *ll = 0;
    if (dl) *dl = 0;
    if ((c->len || c->g4 || c->g5) && ll && dl) {
        *ll = c->len && !par->only_cross_correlations && par->flatsky_nonlocal;
        *dl = c->len && c->den && par->flatsky_local_nonlocal;
    }
    if (!par->sep || !par->multipole_values) return COFFE_EINVAL;
    int st = spline_ops(ops);
    if (st != COFFE_OK) return st;
    return local_setup(par, bg, integral, lc);
}

int coffe_multipoles_init(coffe_parameters_t *par, coffe_background_t *bg, coffe_integral_array_t *integral,
                          coffe_multipoles_array_t *mp)
{
    if (!par || !mp) return done(COFFE_EINVAL);
    coffe_multipoles_free(mp);
    const size_t nz = par->z_mean_len, nl = par->multipole_values_len, ns = par->sep_len;
    const size_t size = nz * nl * ns;
    if (!size) return done(COFFE_OK);
    if (!par->z_mean || nz > INT32_MAX || ns > INT32_MAX || nl > INT32_MAX) return done(COFFE_EINVAL);
    coffe_spline_ops ops;
    local_ctx lc;
    int want_ll = 0, want_dl = 0;
    int st = multipole_setup(par, bg, integral, &ops, &lc, &want_ll, &want_dl);
    if (st != COFFE_OK) return done(st);
    coffe_multipoles_t *arr = (coffe_multipoles_t *)malloc(sizeof(coffe_multipoles_t) * size);
    double *vals = (double *)malloc(sizeof(double) * size);
    double *lens = (want_ll || want_dl) ? (double *)malloc(sizeof(double) * size) : NULL;
    if (!arr || !vals || ((want_ll || want_dl) && !lens)) {
        free(arr);
        free(vals);
        free(lens);
        return done(COFFE_ENOMEM);
    }
    st = local_eval_ls(&lc, &ops, 0, (int)nz, par->z_mean, (int)nl, par->multipole_values, (int)ns, par->sep, vals);
    /* the flat-sky lensing terms and the full-sky integrated ones, as multipoles.c adds
       the integrated outputs */
    if (st == COFFE_OK && lens) {
        st = flat_eval(par, bg, integral, 0, want_ll, want_dl, (int)nz, par->z_mean, (int)nl, par->multipole_values,
                       NULL, (int)ns, par->sep, lens);
        for (size_t q = 0; st == COFFE_OK && q < size; ++q) vals[q] += lens[q];
    }
    free(lens);
    int any = 0;
    const unsigned f = intg_flags(par, &any);
    if (st == COFFE_OK && any && (par->correlation_contrib.len || par->correlation_contrib.g4 ||
                                  par->correlation_contrib.g5)) {
        double *part = (double *)malloc(sizeof(double) * size);
        if (!part) st = COFFE_ENOMEM;
        if (st == COFFE_OK)
            st = intg_multipoles(par, bg, integral, f, -1, (int)nz, par->z_mean, (int)nl, par->multipole_values,
                                 (int)ns, par->sep, part);
        for (size_t q = 0; st == COFFE_OK && q < size; ++q) vals[q] += part[q];
        free(part);
    }
    if (st != COFFE_OK) {
        free(arr);
        free(vals);
        return done(st);
    }
    /* z-major, then l, then separation (multipoles.c:76-86) */
    size_t i = 0;
    for (size_t z = 0; z < nz; ++z)
        for (size_t l = 0; l < nl; ++l)
            for (size_t r = 0; r < ns; ++r, ++i) {
                arr[i].coords.z_mean = par->z_mean[z];
                arr[i].coords.l = par->multipole_values[l];
                arr[i].coords.separation = par->sep[r];
                arr[i].value = vals[i];
            }
    free(vals);
    mp->array = arr;
    mp->size = size;
    return done(COFFE_OK);
}

int coffe_average_multipoles_init(coffe_parameters_t *par, coffe_background_t *bg, coffe_integral_array_t *integral,
                                  coffe_average_multipoles_array_t *ramp)
{
    if (!par || !ramp) return done(COFFE_EINVAL);
    coffe_average_multipoles_free(ramp);
    const size_t nb = par->zmin_len, nl = par->multipole_values_len, ns = par->sep_len;
    const size_t size = nb * nl * ns;
    if (!size) return done(COFFE_OK);
    if (!par->zmin || !par->zmax || par->zmax_len < nb || nl > INT32_MAX || ns > INT32_MAX)
        return done(COFFE_EINVAL);
    coffe_spline_ops ops;
    local_ctx lc;
    int want_ll = 0, want_dl = 0;
    int st = multipole_setup(par, bg, integral, &ops, &lc, &want_ll, &want_dl);
    if (st == COFFE_OK && !bg) st = COFFE_EINVAL;
    if (st != COFFE_OK) return done(st);
    const coffe_correlation_contributions *c = &par->correlation_contrib;
    const int intg = c->len || c->g4 || c->g5;
    coffe_average_multipoles_t *arr = (coffe_average_multipoles_t *)malloc(sizeof(*arr) * size);
    double *vals = (double *)malloc(sizeof(double) * nl * ns);
    double *part = intg ? (double *)malloc(sizeof(double) * nl * ns) : NULL;
    if (!arr || !vals || (intg && !part)) st = COFFE_ENOMEM;
    /* bin-major, then l, then separation (average_multipoles.c:90-104) */
    size_t i = 0;
    for (size_t b = 0; b < nb && st == COFFE_OK; ++b) {
        st = average_eval(&lc, &ops, par, bg, par->zmin[b], par->zmax[b], (int)ns, par->sep, (int)nl,
                          par->multipole_values, vals);
        if (st == COFFE_OK && intg) {
            st = intg_average(par, bg, integral, &ops, -1, want_ll, want_dl, par->zmin[b], par->zmax[b], (int)ns,
                              par->sep, (int)nl, par->multipole_values, part);
            for (size_t q = 0; st == COFFE_OK && q < nl * ns; ++q) vals[q] += part[q];
        }
        for (size_t l = 0; l < nl && st == COFFE_OK; ++l)
            for (size_t r = 0; r < ns; ++r, ++i) {
                arr[i].coords.z_min = par->zmin[b];
                arr[i].coords.z_max = par->zmax[b];
                arr[i].coords.l = par->multipole_values[l];
                arr[i].coords.separation = par->sep[r];
                arr[i].value = vals[l * ns + r];
            }
    }
    free(vals);
    free(part);
    if (st != COFFE_OK) {
        free(arr);
        return done(st);
    }
    ramp->array = arr;
    ramp->size = size;
    return done(COFFE_OK);
}

/* ---- covariance.c -------------------------------------------------------------------- */

double covariance_coefficient(double bias1, double bias2, double growth_rate, int ell)
{
    const double f = growth_rate;
    if (ell == 0) return bias1 * bias2 + (bias1 + bias2) * f / 3.0 + f * f / 5.0;
    if (ell == 2) return 2.0 * (bias1 + bias2) * f / 3.0 + 4.0 * f * f / 7.0;
    if (ell == 4) return 8.0 * f * f / 35.0;
    return 0.0;
}

int coffe_covariance_free(coffe_covariance_array_t *cov)
{
    if (!cov) return done(COFFE_EINVAL);
    free(cov->array);
    cov->array = NULL;
    cov->size = 0;
    return done(COFFE_OK);
}

int coffe_average_covariance_free(coffe_average_covariance_array_t *cov)
{
    if (!cov) return done(COFFE_EINVAL);
    free(cov->array);
    cov->array = NULL;
    cov->size = 0;
    return done(COFFE_OK);
}

/* 1 / (H chi^2 (1 + z)) with H the conformal Hubble rate (covariance.c:150-162) */
static double volume_integrand(const coffe_background_t *bg, const coffe_spline_ops *ops, double z)
{
    const double chi = ops->interp_spline(&bg->comoving_distance, z);
    return 1.0 / ops->interp_spline(&bg->conformal_Hz, z) / (chi * chi) / (1.0 + z);
}

/* int_a^b of the volume integrand: adaptive 61-point Gauss-Kronrod bisection to a
   relative 1e-12 (the integrand is smooth and positive) */
static double volume_integral(const coffe_background_t *bg, const coffe_spline_ops *ops, double a, double b)
{
    double lo[64], hi[64];
    int top = 0;
    double total = 0.0;
    lo[top] = a;
    hi[top] = b;
    ++top;
    while (top > 0) {
        --top;
        const double x0 = lo[top], x1 = hi[top], c = 0.5 * (x0 + x1), h = 0.5 * (x1 - x0);
        double k = GK61K_W[30] * volume_integrand(bg, ops, c), g = 0.0;
        for (int j = 0; j < 30; ++j) {
            const double d = h * GK61K_X[j];
            const double s = volume_integrand(bg, ops, c - d) + volume_integrand(bg, ops, c + d);
            k += GK61K_W[j] * s;
            if (j & 1) g += GK61G_W[j >> 1] * s;
        }
        k *= h;
        g *= h;
        if (fabs(k - g) <= 1e-12 * fabs(k) || top >= 62 || h < 1e-9 * fabs(b - a)) {
            total += k;
        } else {
            lo[top] = c;
            hi[top] = x1;
            ++top;
            lo[top] = x0;
            hi[top] = c;
            ++top;
        }
    }
    return total;
}

int coffe_covariance_init(coffe_parameters_t *par, coffe_background_t *bg, coffe_covariance_array_t *cov_mp,
                          coffe_average_covariance_array_t *cov_ramp)
{
    if (!par) return done(COFFE_EINVAL);
    const int avg = par->output_type == COFFE_REF_COVARIANCE_AVERAGE_MULTIPOLES;
    if (par->output_type != COFFE_REF_COVARIANCE_MULTIPOLES && !avg) return done(COFFE_OK);
    if (!bg || !cov_mp || !cov_ramp) return done(COFFE_EINVAL);
    coffe_covariance_free(cov_mp);
    coffe_average_covariance_free(cov_ramp);
    /* One method here: the STANDARD one (the reference's default) is computed by FFTLog
     * too; the reference's tests hold both to the same fixtures (rtol 5e-4). */
    if ((par->covariance_integration_method != COFFE_REF_COVARIANCE_FFTLOG &&
         par->covariance_integration_method != COFFE_REF_COVARIANCE_STANDARD) ||
        par->covariance_interpolation_method != COFFE_REF_INTERP2D_BICUBIC || par->pk_type != 0)
        return done(COFFE_EUNSUPPORTED);
    const size_t nbin = avg ? par->zmin_len : par->z_mean_len;
    const size_t nl = par->multipole_values_len, R = par->sep_len;
    if (!nbin || !nl || !R) return done(COFFE_OK);
    /* the per-bin parameters (covariance.c:788-860 reads them for every bin) */
    if (par->density1_len < nbin || par->density2_len < nbin || par->fsky_len < nbin || par->pixelsize_len < nbin ||
        !par->density1 || !par->density2 || !par->fsky || !par->pixelsize || !par->sep || !par->multipole_values ||
        (avg ? (!par->zmin || !par->zmax || par->zmax_len < nbin)
             : (!par->z_mean || !par->deltaz || par->deltaz_len < nbin)) ||
        nl > 64 || R > INT32_MAX || nbin > INT32_MAX)
        return done(COFFE_EINVAL);
    coffe_spline_ops ops;
    int st = spline_ops(&ops);
    if (st != COFFE_OK) return done(st);
    nodes pk;
    if (nodes_of(&par->power_spectrum_norm, &pk) != COFFE_OK) return done(COFFE_EINVAL);
    coffe_engine *e = coffe_engine_default();
    if (!e) return done(COFFE_ENODEVICE);
    const coffe_correlation_contributions *c = &par->correlation_contrib;
    double *bias = (double *)calloc(4 * nbin, sizeof(double));
    double *f = (double *)calloc(nbin, sizeof(double)), *D1 = (double *)calloc(nbin, sizeof(double));
    double *vol = (double *)calloc(nbin, sizeof(double));
    const size_t nout = nbin * nl * R * nl * R;
    double *h_out = (double *)malloc(sizeof(double) * nout);
    if (!bias || !f || !D1 || !vol || !h_out) st = COFFE_ENOMEM;
    const coffe_interpolation *gb[4] = {&par->galaxy_bias1, &par->galaxy_bias2, &par->galaxy_bias3,
                                        &par->galaxy_bias4};
    const double hH0 = REF_H0 * par->h;
    for (size_t k = 0; k < nbin && st == COFFE_OK; ++k) {
        double z;
        if (avg) {
            z = 0.5 * (par->zmin[k] + par->zmax[k]);
            /* covariance.c:808-837, the factor (H0 h)^-2 included */
            vol[k] = 4.0 * M_PI * par->fsky[k] / volume_integral(bg, &ops, par->zmin[k], par->zmax[k]) / (hH0 * hH0);
        } else {
            z = par->z_mean[k];
            const double c0 = ops.interp_spline(&bg->comoving_distance, z - par->deltaz[k]);
            const double c1 = ops.interp_spline(&bg->comoving_distance, z + par->deltaz[k]);
            vol[k] = 4.0 * M_PI * par->fsky[k] * (c1 * c1 * c1 - c0 * c0 * c0) / 3.0;
        }
        /* biases only with a density contribution, f only with RSD (covariance.c:842-860) */
        if (c->den)
            for (int i = 0; i < 4; ++i) bias[4 * k + i] = ops.interp_spline(gb[i], z);
        if (c->rsd) f[k] = ops.interp_spline(&bg->f, z);
        D1[k] = ops.interp_spline(&bg->D1, z);
    }
    const int pops[4] = {par->covariance_pop1, par->covariance_pop2, par->covariance_pop3, par->covariance_pop4};
    coffe_cov_plan *plan = NULL;
    void *d_k = NULL, *d_pk = NULL, *d_integ = NULL, *d_out = NULL;
    if (st == COFFE_OK)
        st = coffe_cov_plan_create(e, par->covariance_integration_bins, par->k_min_norm, par->k_max_norm, (int)nl,
                                   par->multipole_values, (int)R, par->sep, &plan);
    if (st == COFFE_OK) {
        hipSetDevice(e->device);
        const size_t ninteg = 2 * (size_t)coffe_cov_plan_pairs(plan) * R * R;
        st = ce_dev_alloc(&d_k, sizeof(double) * pk.n);
        if (st == COFFE_OK) st = ce_dev_alloc(&d_pk, sizeof(double) * pk.n);
        if (st == COFFE_OK) st = ce_dev_alloc(&d_integ, sizeof(double) * ninteg);
        if (st == COFFE_OK) st = ce_dev_alloc(&d_out, sizeof(double) * nout);
    }
    if (st == COFFE_OK) st = ce_hip_ok(hipMemcpy(d_k, pk.x, sizeof(double) * pk.n, hipMemcpyHostToDevice));
    if (st == COFFE_OK) st = ce_hip_ok(hipMemcpy(d_pk, pk.y, sizeof(double) * pk.n, hipMemcpyHostToDevice));
    const int64_t n_in = (int64_t)pk.n;
    if (st == COFFE_OK) st = coffe_cov_integrals(plan, 1, &n_in, (const double *)d_k, (const double *)d_pk,
                                                 (double *)d_integ, NULL);
    if (st == COFFE_OK)
        st = coffe_cov_assemble(plan, 1, (int)nbin, pops, bias, f, D1, par->density1, par->density2, vol,
                                par->pixelsize, par->covariance_poisson, par->covariance_mixed,
                                par->covariance_cosmic, par->covariance_window, (const double *)d_integ,
                                (double *)d_out, NULL);
    if (st == COFFE_OK) st = coffe_engine_synchronize(e);
    if (st == COFFE_OK) st = ce_hip_ok(hipMemcpy(h_out, d_out, sizeof(double) * nout, hipMemcpyDeviceToHost));
    if (d_k) hipFree(d_k);
    if (d_pk) hipFree(d_pk);
    if (d_integ) hipFree(d_integ);
    if (d_out) hipFree(d_out);
    if (plan) coffe_cov_plan_destroy(plan);
    free(bias);
    free(f);
    free(D1);
    free(vol);
    if (st == COFFE_OK) {
        /* records z, l1, r1, l2, r2 (covariance.c:1043-1074), the same order as d_out */
        if (avg) {
            coffe_average_covariance_t *arr = (coffe_average_covariance_t *)malloc(sizeof(*arr) * nout);
            if (!arr) st = COFFE_ENOMEM;
            size_t i = 0;
            for (size_t k = 0; k < nbin && arr; ++k)
                for (size_t a = 0; a < nl; ++a)
                    for (size_t m = 0; m < R; ++m)
                        for (size_t b = 0; b < nl; ++b)
                            for (size_t n = 0; n < R; ++n, ++i) {
                                arr[i].coords.z_min = par->zmin[k];
                                arr[i].coords.z_max = par->zmax[k];
                                arr[i].coords.l1 = par->multipole_values[a];
                                arr[i].coords.l2 = par->multipole_values[b];
                                arr[i].coords.separation1 = par->sep[m];
                                arr[i].coords.separation2 = par->sep[n];
                                arr[i].value = h_out[i];
                            }
            if (arr) {
                cov_ramp->array = arr;
                cov_ramp->size = nout;
            }
        } else {
            coffe_covariance_t *arr = (coffe_covariance_t *)malloc(sizeof(*arr) * nout);
            if (!arr) st = COFFE_ENOMEM;
            size_t i = 0;
            for (size_t k = 0; k < nbin && arr; ++k)
                for (size_t a = 0; a < nl; ++a)
                    for (size_t m = 0; m < R; ++m)
                        for (size_t b = 0; b < nl; ++b)
                            for (size_t n = 0; n < R; ++n, ++i) {
                                arr[i].coords.z_mean = par->z_mean[k];
                                arr[i].coords.l1 = par->multipole_values[a];
                                arr[i].coords.l2 = par->multipole_values[b];
                                arr[i].coords.separation1 = par->sep[m];
                                arr[i].coords.separation2 = par->sep[n];
                                arr[i].value = h_out[i];
                            }
            if (arr) {
                cov_mp->array = arr;
                cov_mp->size = nout;
            }
        }
    }
    free(h_out);
    return done(st);
}

/* |a - b| / |a| < rel when both are non-zero, else |a - b| < abs (common.c:1206-1219) */
static int approx_equal(double a, double b, double rel, double abs_eps)
{
    if (a != 0.0 && b != 0.0) return fabs(a - b) / fabs(a) < rel;
    return fabs(a - b) < abs_eps;
}

coffe_covariance_t coffe_covariance_find(const coffe_covariance_array_t *cov, double z_mean, int l1, int l2,
                                         double sep1, double sep2)
{
    if (cov && cov->array)
        for (size_t i = 0; i < cov->size; ++i) {
            const coffe_covariance_t *r = &cov->array[i];
            if (r->coords.l1 == l1 && r->coords.l2 == l2 && approx_equal(r->coords.separation1, sep1, 1e-3, 0.0) &&
                approx_equal(r->coords.separation2, sep2, 1e-3, 0.0) &&
                approx_equal(r->coords.z_mean, z_mean, 1e-3, 0.0)) {
                done(COFFE_OK);
                return *r;
            }
        }
    coffe_covariance_t none;
    memset(&none, 0, sizeof(none));
    none.coords.l1 = -1;
    none.coords.l2 = -1;
    none.value = NAN;
    done(COFFE_EDOMAIN);
    return none;
}
