/*
 * Host side of the MI355X spherical-Bessel integral engine (C99).
 *
 * Owns the per-device state (stream, cache of spectrum-independent grid tables),
 * validates inputs the way the reference would have exit()ed on them, builds the
 * work lists of a batch and launches the HIP kernels through the launchers of
 * ce_internal.h.  Reference behaviour restated here:
 *   - q selection, twofast_selectqnu (twofast.c:240-250);
 *   - the output grid r_i = r0 (kmax/kmin)^(i/N) (twofast.c:415-417) and the
 *     table layout [r = 0, 42 small separations, N FFT nodes]
 *     (integrals.c:474, 545-567), whose Akima lookups at a separation touch
 *     nodes j-2..j+3 of the interval j containing it (common.c:704-710);
 *   - window-limit validation (twofast.c:62, 81-88) and non-finite checks
 *     (twofast.c:122, 210) turned into return codes.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "coffe_engine.h"
#include "coffe_integrals.h"
#include "ce_internal.h"
#include "ce_host.h"

#define CE_THREADS_DEFAULT 512

static __thread int ce_last_error = 0;

int coffe_engine_last_error(void) { return ce_last_error; }

void ce_set_last_error(int code) { ce_last_error = code; }

const char *coffe_engine_strerror(int code)
{
    switch (code) {
    case COFFE_OK: return "success";
    case COFFE_EINVAL: return "invalid argument";
    case COFFE_ENOMEM: return "out of memory";
    case COFFE_EDEVICE: return "HIP device error";
    case COFFE_EUNSUPPORTED: return "unsupported size or case (an odd or too long FFT length, a batch length "
                                    "outside the LDS transforms, or an unsupported option)";
    case COFFE_EDOMAIN: return "argument outside the domain (window limits or separation outside the table)";
    case COFFE_ENONFINITE: return "non-finite value in input or kernel";
    case COFFE_ENODEVICE: return "no HIP device available";
    case COFFE_ENOCONV: return "adaptive quadrature did not converge";
    default: return "unknown error";
    }
}

/* ---- grids ------------------------------------------------------------------ */

typedef struct ce_grid {
    struct ce_grid *next;
    /* key */
    int N;
    double kmin, kmax, k0, r0;
    int ntab;
    ce_table_desc tabs[CE_MAX_TABLES];
    int nq;
    double qprime[2];
    /* host */
    double *h_p;
    /* device */
    double *d_p, *d_tw, *d_bw[2], *d_U[CE_MAX_TABLES], *d_os[CE_MAX_TABLES];
    ce_radix_plan rp;   /* npass 0: no LDS schedule (N > 16384 or N/2 not 2^a 3^b 5^c) */
    ce_anyfft *fft;     /* the length-N transform of the general path, made on first use */
} ce_grid;

static int hip_ok(hipError_t e) { return e == hipSuccess ? COFFE_OK : COFFE_EDEVICE; }

int ce_hip_ok(hipError_t e) { return hip_ok(e); }

/* allocation-failure injection (coffe_engine_inject_alloc_failure): the countdown of
   device allocations until the one that fails, 0 = off */
static int ce_fail_countdown = 0;

void coffe_engine_inject_alloc_failure(int n) { __atomic_store_n(&ce_fail_countdown, n > 0 ? n : 0, __ATOMIC_SEQ_CST); }

static int injected_failure(void)
{
    int c = __atomic_load_n(&ce_fail_countdown, __ATOMIC_SEQ_CST);
    while (c > 0) {
        if (__atomic_compare_exchange_n(&ce_fail_countdown, &c, c - 1, 0, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST))
            return c == 1;
    }
    return 0;
}

static int dev_alloc(void **p, size_t bytes)
{
    if (bytes == 0) bytes = 8;
    if (injected_failure()) { *p = NULL; return COFFE_ENOMEM; }
    return hipMalloc(p, bytes) == hipSuccess ? COFFE_OK : COFFE_ENOMEM;
}

int ce_dev_alloc(void **p, size_t bytes) { return dev_alloc(p, bytes); }

void ce_twiddles(int N, double *out)
{
    for (int k = 0; k < N; ++k) {
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)N;
        out[2 * k] = (double)cosl(a);
        out[2 * k + 1] = (double)sinl(a);
    }
}

double ce_selectqnu(double l, double nu)
{
    const double n1 = 0.9, n2 = 0.9999;
    const double qmin = fmax(n2 - 1.0 - nu, -l);
    const double qmax = fmin(n1 + 3.0 - nu, 2.0);
    double q = (2. + n1 + n2) / 2. - nu;
    if (!(qmin < q && q < qmax)) q = (qmin + 2. * qmax) / 3.;
    return q;
}

/* the schedule packed 3 bits a pass, first pass lowest (ce_radix_plan.code) */
static void radix_code(ce_radix_plan *rp)
{
    rp->code = 0;
    for (int p = rp->npass - 1; p >= 0; --p) {
        const int r = rp->radix[p];
        const uint32_t c = r == 2 ? 1 : r == 3 ? 2 : r == 4 ? 3 : r == 5 ? 4 : r == 8 ? 5 : 7;
        rp->code = (rp->code << 3) | c;
    }
}

int ce_radix_schedule(int M, ce_radix_plan *rp) { return ce_radix_schedule_ex(M, rp, 0); }

int ce_radix_schedule_ex(int M, ce_radix_plan *rp, int pair10)
{
    rp->npass = 0;
    int m = M;
    int n10 = 0;
    if (pair10 && M > 0) {
        int a = 0, b = 0, t = M;
        while (t % 2 == 0) { ++a; t /= 2; }
        t = M;
        while (t % 5 == 0) { ++b; t /= 5; }
        n10 = a < b ? a : b;
        for (int i = 0; i < n10; ++i) m /= 10;
    }
    /* the power of two as radix-8 passes, then one radix-4 pass for a remaining 4, or
       two radix-4 passes in place of 8 x 2 for a remaining 2 (same pass count; the
       radix-2 pass with four butterflies per thread measured slower) */
    int a = 0;
    while (m % 2 == 0) { ++a; m /= 2; }
    int n8 = a / 3;
    const int rem = a % 3;
    if (rem == 1 && n8 > 0) --n8;
    for (int i = 0; i < n8; ++i) rp->radix[rp->npass++] = 8;
    if (rem == 2) rp->radix[rp->npass++] = 4;
    if (rem == 1) {
        if (a > 1) { rp->radix[rp->npass++] = 4; rp->radix[rp->npass++] = 4; }
        else rp->radix[rp->npass++] = 2;
    }
    while (m % 5 == 0) { if (rp->npass >= CE_MAX_PASSES) return COFFE_EUNSUPPORTED; rp->radix[rp->npass++] = 5; m /= 5; }
    while (m % 3 == 0) { if (rp->npass >= CE_MAX_PASSES) return COFFE_EUNSUPPORTED; rp->radix[rp->npass++] = 3; m /= 3; }
    for (int i = 0; i < n10; ++i) {
        if (rp->npass >= CE_MAX_PASSES) return COFFE_EUNSUPPORTED;
        rp->radix[rp->npass++] = 10;
    }
    if (m != 1 || rp->npass > 10) return COFFE_EUNSUPPORTED;
    radix_code(rp);
    return COFFE_OK;
}

static void grid_free(ce_grid *g)
{
    if (!g) return;
    free(g->h_p);
    hipFree(g->d_p);
    hipFree(g->d_tw);
    for (int q = 0; q < 2; ++q) hipFree(g->d_bw[q]);
    for (int t = 0; t < CE_MAX_TABLES; ++t) { hipFree(g->d_U[t]); hipFree(g->d_os[t]); }
    ce_anyfft_destroy(g->fft);
    free(g);
}

static int same_tabs(const ce_grid *g, int ntab, const ce_table_desc *tabs)
{
    if (g->ntab != ntab) return 0;
    for (int t = 0; t < ntab; ++t)
        if (g->tabs[t].n != tabs[t].n || g->tabs[t].l != tabs[t].l || g->tabs[t].rpow != tabs[t].rpow) return 0;
    return 1;
}

/* assign q' = q(l, nu) + nu classes; at most two distinct values */
static int assign_q(int ntab, ce_table_desc *tabs, double *qnu, int *nq, double *qprime)
{
    *nq = 0;
    for (int t = 0; t < ntab; ++t) {
        qnu[t] = ce_selectqnu(tabs[t].l, tabs[t].n);
        const double qp = qnu[t] + tabs[t].n;
        int found = -1;
        for (int q = 0; q < *nq; ++q)
            if (qprime[q] == qp) found = q;
        if (found < 0) {
            if (*nq >= 2) return COFFE_EUNSUPPORTED;
            qprime[*nq] = qp;
            found = (*nq)++;
        }
        tabs[t].qidx = found;
    }
    return COFFE_OK;
}

/* grid tables for (N, kmin, kmax, k0, r0, tables); cached on the engine */
static int grid_get(coffe_engine *e, int N, double kmin, double kmax, double k0, double r0, int ntab,
                    const ce_table_desc *tabs_in, ce_grid **out)
{
    if (N < 4 || (N & 1) || N > CE_ANY_MAX_N) return COFFE_EUNSUPPORTED;
    if (!(kmin > 0) || !(kmax > kmin) || !(k0 > 0) || !(r0 > 0) || !isfinite(kmax)) return COFFE_EINVAL;
    if (ntab < 1 || ntab > CE_MAX_TABLES) return COFFE_EINVAL;
    pthread_mutex_lock(&e->lock);
    for (ce_grid *g = e->grids; g; g = g->next) {
        if (g->N == N && g->kmin == kmin && g->kmax == kmax && g->k0 == k0 && g->r0 == r0 &&
            same_tabs(g, ntab, tabs_in)) {
            pthread_mutex_unlock(&e->lock);
            *out = g;
            return COFFE_OK;
        }
    }
    pthread_mutex_unlock(&e->lock);
    ce_grid *g = (ce_grid *)calloc(1, sizeof(*g));
    if (!g) return COFFE_ENOMEM;
    g->N = N; g->kmin = kmin; g->kmax = kmax; g->k0 = k0; g->r0 = r0; g->ntab = ntab;
    memcpy(g->tabs, tabs_in, sizeof(ce_table_desc) * (size_t)ntab);
    double qnu[CE_MAX_TABLES];
    int st = assign_q(ntab, g->tabs, qnu, &g->nq, g->qprime);
    /* the fused LDS kernels' schedule; lengths without one take the general path */
    if (st == COFFE_OK && (N > 16384 || ce_radix_schedule(N / 2, &g->rp) != COFFE_OK)) memset(&g->rp, 0, sizeof(g->rp));
    /* window limits of twofast_fft_input (twofast.c:283-289) */
    const double xmax = k0 * pow(kmax / kmin, (double)(N - 1) / N);
    const double xl = exp(0.46) * kmin, xr = exp(-0.46) * xmax;
    if (st == COFFE_OK && !(kmin <= xl && xl <= xr && xr <= xmax)) st = COFFE_EDOMAIN;
    g->h_p = (double *)malloc(sizeof(double) * (size_t)N);
    double *h_tw = (double *)malloc(sizeof(double) * 2 * (size_t)N);
    if (st == COFFE_OK && (!g->h_p || !h_tw)) st = COFFE_ENOMEM;
    if (st == COFFE_OK) {
        for (int i = 0; i < N; ++i) g->h_p[i] = pow(kmax / kmin, (double)i / N);
        ce_twiddles(N, h_tw);
    }
    const size_t nb = sizeof(double) * (size_t)N;
    const size_t cb = sizeof(double) * 2 * (size_t)(N / 2 + 1);
    if (st == COFFE_OK) st = dev_alloc((void **)&g->d_p, nb);
    if (st == COFFE_OK) st = dev_alloc((void **)&g->d_tw, 2 * nb);
    for (int q = 0; q < g->nq && st == COFFE_OK; ++q) st = dev_alloc((void **)&g->d_bw[q], nb);
    for (int t = 0; t < ntab && st == COFFE_OK; ++t) {
        st = dev_alloc((void **)&g->d_U[t], cb);
        if (st == COFFE_OK) st = dev_alloc((void **)&g->d_os[t], nb);
    }
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(g->d_p, g->h_p, nb, hipMemcpyHostToDevice, e->stream));
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(g->d_tw, h_tw, 2 * nb, hipMemcpyHostToDevice, e->stream));
    if (st == COFFE_OK) {
        st = ce_launch_grid_tables(e->stream, N, kmin, kmax, k0, r0, g->d_p, ntab, g->tabs, qnu, g->nq, g->qprime,
                                   g->d_U, g->d_os, g->d_bw) ? COFFE_EDEVICE : COFFE_OK;
    }
    if (st == COFFE_OK) st = hip_ok(hipStreamSynchronize(e->stream));
    free(h_tw);
    if (st == COFFE_OK) {
        /* the gamma-ratio kernel must be finite (twofast.c:210-223) */
        double *hU = (double *)malloc(cb);
        if (!hU) st = COFFE_ENOMEM;
        for (int t = 0; t < ntab && st == COFFE_OK; ++t) {
            st = hip_ok(hipMemcpy(hU, g->d_U[t], cb, hipMemcpyDeviceToHost));
            for (int k = 0; k < N + 2 && st == COFFE_OK; ++k)
                if (!isfinite(hU[k])) st = COFFE_ENONFINITE;
        }
        free(hU);
    }
    if (st != COFFE_OK) { grid_free(g); return st; }
    pthread_mutex_lock(&e->lock);
    g->next = e->grids;
    e->grids = g;
    pthread_mutex_unlock(&e->lock);
    *out = g;
    return COFFE_OK;
}

static void grid_to_dev(const ce_grid *g, ce_grid_dev *d)
{
    memset(d, 0, sizeof(*d));
    d->rp = g->rp;
    d->N = g->N;
    d->M = g->N / 2;
    d->kmin = g->kmin;
    d->kmax = g->kmax;
    d->G = log(g->kmax / g->kmin);
    d->k0 = g->k0;
    d->r0 = g->r0;
    d->p = g->d_p;
    d->tw = g->d_tw;
    for (int q = 0; q < 2; ++q) d->bw[q] = g->d_bw[q];
    for (int t = 0; t < CE_MAX_TABLES; ++t) { d->U[t] = g->d_U[t]; d->os[t] = g->d_os[t]; }
}

/* ---- engine ----------------------------------------------------------------- */

int coffe_engine_create(int device, coffe_engine **out)
{
    if (!out) return COFFE_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return COFFE_ENODEVICE;
    if (device < 0 || device >= ndev) return COFFE_EINVAL;
    coffe_engine *e = (coffe_engine *)calloc(1, sizeof(*e));
    if (!e) return COFFE_ENOMEM;
    e->device = device;
    e->threads = CE_THREADS_DEFAULT;
    const char *t = getenv("COFFE_ENGINE_THREADS");
    if (t) {
        int v = atoi(t);
        if (v == 256 || v == 512 || v == 1024) e->threads = v;
    }
    pthread_mutex_init(&e->lock, NULL);
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&e->stream) != hipSuccess) {
        free(e);
        return COFFE_EDEVICE;
    }
    if (hipEventCreateWithFlags(&e->ev_scratch, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev_stage[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev_stage[1], hipEventDisableTiming) != hipSuccess) {
        hipStreamDestroy(e->stream);
        free(e);
        return COFFE_EDEVICE;
    }
    *out = e;
    return COFFE_OK;
}

void coffe_engine_destroy(coffe_engine *e)
{
    if (!e) return;
    hipStreamSynchronize(e->stream);
    if (e->scratch_pending) hipEventSynchronize(e->ev_scratch);
    for (int i = 0; i < 2; ++i)
        if (e->stage_pending[i]) hipEventSynchronize(e->ev_stage[i]);
    ce_grid *g = e->grids;
    while (g) { ce_grid *n = g->next; grid_free(g); g = n; }
    hipFree(e->d_scratch);
    for (int i = 0; i < 2; ++i) {
        hipHostFree(e->h_pinned[i]);
        hipEventDestroy(e->ev_stage[i]);
    }
    hipEventDestroy(e->ev_scratch);
    hipStreamDestroy(e->stream);
    pthread_mutex_destroy(&e->lock);
    free(e);
}

int coffe_engine_synchronize(coffe_engine *e)
{
    if (!e) return COFFE_EINVAL;
    return hip_ok(hipStreamSynchronize(e->stream));
}

static coffe_engine *g_default = NULL;
static pthread_mutex_t g_default_lock = PTHREAD_MUTEX_INITIALIZER;

coffe_engine *coffe_engine_default(void)
{
    pthread_mutex_lock(&g_default_lock);
    if (!g_default) {
        coffe_engine *e = NULL;
        if (coffe_engine_create(0, &e) == COFFE_OK) g_default = e;
    }
    pthread_mutex_unlock(&g_default_lock);
    return g_default;
}

/* ---- batch ------------------------------------------------------------------ */

struct coffe_batch {
    coffe_engine *e;
    int nspec, ntab, nsep, nz, nl, nq, nunit, ngrid, nchunk;
    int max_M, max_N, threads;
    int64_t total_in;
    ce_grid **grids;
    ce_spec_dev *h_spec;
    /* device */
    ce_spec_dev *d_spec;
    ce_grid_dev *d_grids;
    ce_table_desc *d_tabs;
    ce_unit_dev *d_units;
    ce_chunk_dev *d_chunks;
    int32_t *d_ranges; /* resample index ranges per chunk (4 each), from the prepare stage */
    int32_t *d_packs;  /* table workgroups: first unit, unit count, groups, 0 */
    int npack, lt_off;
    int32_t *d_sidx;
    double *d_sep;
    double *d_ws, *d_xf, *d_xb, *d_wbuf, *d_itab;
    int *d_srstat;     /* status of the small-r integrals, allocated on first use */
};

int64_t coffe_batch_input_size(const coffe_batch *b) { return b ? b->total_in : 0; }

void coffe_batch_destroy(coffe_batch *b)
{
    if (!b) return;
    free(b->grids);
    free(b->h_spec);
    hipFree(b->d_spec);
    hipFree(b->d_grids);
    hipFree(b->d_tabs);
    hipFree(b->d_units);
    hipFree(b->d_chunks);
    hipFree(b->d_ranges);
    hipFree(b->d_packs);
    hipFree(b->d_sidx);
    hipFree(b->d_sep);
    hipFree(b->d_ws);
    hipFree(b->d_xf);
    hipFree(b->d_xb);
    hipFree(b->d_wbuf);
    hipFree(b->d_itab);
    hipFree(b->d_srstat);
    free(b);
}

/* interval j of the FFT nodes r_i = r0 p_i containing sep, with x_j <= sep < x_{j+1} */
static int sep_interval(const ce_grid *g, double sep, int32_t *j_out)
{
    const int N = g->N;
    const double r_lo = g->r0 * g->h_p[0];
    if (!(sep >= r_lo) || !(sep <= g->r0 * g->h_p[N - 1])) return COFFE_EDOMAIN;
    int lo = 0, hi = N - 1;
    while (hi > lo + 1) {
        const int m = (lo + hi) / 2;
        if (g->r0 * g->h_p[m] > sep) hi = m;
        else lo = m;
    }
    /* the Akima stencil must stay among the FFT nodes (no small-r nodes, no
       extrapolated end slopes) */
    if (lo < 2 || lo + 3 > N - 1) return COFFE_EDOMAIN;
    *j_out = lo;
    return COFFE_OK;
}

typedef struct { int idx; int N; } unit_sort;
static int cmp_unit(const void *a, const void *b)
{
    const unit_sort *x = (const unit_sort *)a, *y = (const unit_sort *)b;
    if (x->N != y->N) return y->N - x->N;
    return x->idx - y->idx;
}

int coffe_batch_create(coffe_engine *e, int nspec, const int64_t *n_in, const int64_t *n_fft, const double *kmin,
                       const double *kmax, int ntab, const int *tab_n, const int *tab_l, int nsep, const double *sep,
                       int nz, int nl, coffe_batch **out)
{
    if (!e || !out || nspec <= 0 || !n_in || !n_fft || !kmin || !kmax || ntab <= 0 || ntab > CE_MAX_TABLES ||
        !tab_n || !tab_l || nsep <= 0 || !sep || nz <= 0 || nl <= 0)
        return COFFE_EINVAL;
    hipSetDevice(e->device);
    coffe_batch *b = (coffe_batch *)calloc(1, sizeof(*b));
    if (!b) return COFFE_ENOMEM;
    b->e = e;
    b->nspec = nspec; b->ntab = ntab; b->nsep = nsep; b->nz = nz; b->nl = nl;
    b->threads = e->threads;
    ce_table_desc tabs[CE_MAX_TABLES];
    for (int t = 0; t < ntab; ++t) {
        if (tab_n[t] < -2 || tab_n[t] > 4 || tab_l[t] < 0 || tab_l[t] > 4) { coffe_batch_destroy(b); return COFFE_EINVAL; }
        tabs[t].n = tab_n[t];
        tabs[t].l = tab_l[t];
        tabs[t].rpow = 1;
        tabs[t].qidx = 0;
    }
    b->grids = (ce_grid **)calloc((size_t)nspec, sizeof(ce_grid *));
    b->h_spec = (ce_spec_dev *)calloc((size_t)nspec, sizeof(ce_spec_dev));
    int32_t *spec_grid = (int32_t *)calloc((size_t)nspec, sizeof(int32_t));
    if (!b->grids || !b->h_spec || !spec_grid) { free(spec_grid); coffe_batch_destroy(b); return COFFE_ENOMEM; }
    int st = COFFE_OK;
    int64_t in_off = 0, xf_off = 0, xb_off = 0, ws_off = 0;
    for (int s = 0; s < nspec && st == COFFE_OK; ++s) {
        if (n_in[s] < 5 || n_fft[s] < 4 || n_fft[s] > 16384) { st = n_in[s] < 5 ? COFFE_EINVAL : COFFE_EUNSUPPORTED; break; }
        ce_grid *g = NULL;
        st = grid_get(e, (int)n_fft[s], kmin[s], kmax[s], kmin[s], 1.0, ntab, tabs, &g);
        if (st == COFFE_OK && g->rp.npass == 0) st = COFFE_EUNSUPPORTED;   /* the batch runs the LDS kernels */
        if (st != COFFE_OK) break;
        int gi = -1;
        for (int q = 0; q < b->ngrid; ++q) if (b->grids[q] == g) gi = q;
        if (gi < 0) { gi = b->ngrid; b->grids[b->ngrid++] = g; }
        spec_grid[s] = gi;
        const int M = g->N / 2;
        if (M > b->max_M) b->max_M = M;
        if (g->N > b->max_N) b->max_N = g->N;
        b->nq = g->nq;
        ce_spec_dev *d = &b->h_spec[s];
        d->in_off = in_off;
        d->n_in = (int32_t)n_in[s];
        d->grid = gi;
        d->xf_off = xf_off;
        d->xb_off = xb_off;
        d->ws_off = ws_off;
        in_off += n_in[s];
        xf_off += (int64_t)g->nq * (M + 1);
        xb_off += (int64_t)ntab * (M + 1);
        ws_off += 2 * (int64_t)g->N;
        b->nchunk += (int)((n_in[s] - 1 + CE_CHUNK - 1) / CE_CHUNK);
    }
    b->total_in = in_off;
    /* every FFT pass needs at least M/8 threads (at most 4 butterflies each) */
    while (st == COFFE_OK && b->threads < b->max_M / 8) b->threads *= 2;
    /* separation intervals per grid */
    int32_t *sidx = NULL;
    if (st == COFFE_OK) {
        sidx = (int32_t *)malloc(sizeof(int32_t) * (size_t)b->ngrid * (size_t)nsep);
        if (!sidx) st = COFFE_ENOMEM;
        for (int gi = 0; gi < b->ngrid && st == COFFE_OK; ++gi)
            for (int r = 0; r < nsep && st == COFFE_OK; ++r) st = sep_interval(b->grids[gi], sep[r], &sidx[gi * nsep + r]);
    }
    /* work units, largest FFT first */
    ce_unit_dev *units = NULL;
    if (st == COFFE_OK) {
        b->nunit = nspec * ntab;
        unit_sort *us = (unit_sort *)malloc(sizeof(unit_sort) * (size_t)b->nunit);
        units = (ce_unit_dev *)malloc(sizeof(ce_unit_dev) * (size_t)b->nunit);
        if (!us || !units) st = COFFE_ENOMEM;
        for (int i = 0; i < b->nunit && st == COFFE_OK; ++i) { us[i].idx = i; us[i].N = (int)n_fft[i / ntab]; }
        if (st == COFFE_OK) {
            qsort(us, (size_t)b->nunit, sizeof(unit_sort), cmp_unit);
            for (int i = 0; i < b->nunit; ++i) { units[i].spec = us[i].idx / ntab; units[i].tab = us[i].idx % ntab; }
        }
        free(us);
    }
    ce_chunk_dev *chunks = NULL;
    if (st == COFFE_OK) {
        chunks = (ce_chunk_dev *)malloc(sizeof(ce_chunk_dev) * (size_t)b->nchunk);
        if (!chunks) st = COFFE_ENOMEM;
        int c = 0;
        for (int s = 0; s < nspec && st == COFFE_OK; ++s) {
            const int nint = (int)n_in[s] - 1;
            for (int a = 0; a < nint; a += CE_CHUNK) {
                chunks[c].spec = s;
                chunks[c].grid = b->h_spec[s].grid;
                chunks[c].in_off = b->h_spec[s].in_off;
                chunks[c].ws_off = b->h_spec[s].ws_off;
                chunks[c].n_in = b->h_spec[s].n_in;
                chunks[c].id = c;
                chunks[c].a = a;
                chunks[c].b = a + CE_CHUNK < nint ? a + CE_CHUNK : nint;
                ++c;
            }
        }
    }
    ce_grid_dev *hg = NULL;
    if (st == COFFE_OK) {
        hg = (ce_grid_dev *)calloc((size_t)b->ngrid, sizeof(ce_grid_dev));
        if (!hg) st = COFFE_ENOMEM;
        for (int gi = 0; gi < b->ngrid && st == COFFE_OK; ++gi) grid_to_dev(b->grids[gi], &hg[gi]);
    }
    if (st == COFFE_OK) {
        /* the q classes are the same for every grid (they depend on the tables only) */
        memcpy(tabs, b->grids[0]->tabs, sizeof(ce_table_desc) * (size_t)ntab);
    }
    if (st == COFFE_OK) st = dev_alloc((void **)&b->d_spec, sizeof(ce_spec_dev) * (size_t)nspec);
    if (st == COFFE_OK) st = dev_alloc((void **)&b->d_grids, sizeof(ce_grid_dev) * (size_t)b->ngrid);
    if (st == COFFE_OK) st = dev_alloc((void **)&b->d_tabs, sizeof(ce_table_desc) * (size_t)ntab);
    if (st == COFFE_OK) st = dev_alloc((void **)&b->d_units, sizeof(ce_unit_dev) * (size_t)b->nunit);
    if (st == COFFE_OK) st = dev_alloc((void **)&b->d_chunks, sizeof(ce_chunk_dev) * (size_t)b->nchunk);
    if (st == COFFE_OK) st = dev_alloc((void **)&b->d_ranges, 4 * sizeof(int32_t) * (size_t)b->nchunk);
    if (st == COFFE_OK) st = hip_ok(hipMemset(b->d_ranges, 0, 4 * sizeof(int32_t) * (size_t)b->nchunk));
    if (st == COFFE_OK) st = dev_alloc((void **)&b->d_sidx, sizeof(int32_t) * (size_t)b->ngrid * (size_t)nsep);
    if (st == COFFE_OK) st = dev_alloc((void **)&b->d_sep, sizeof(double) * (size_t)nsep);
    if (st == COFFE_OK) st = dev_alloc((void **)&b->d_ws, sizeof(double) * (size_t)(ws_off > 0 ? ws_off : 1));
    if (st == COFFE_OK) st = dev_alloc((void **)&b->d_xf, 2 * sizeof(double) * (size_t)xf_off);
    if (st == COFFE_OK) st = dev_alloc((void **)&b->d_xb, 2 * sizeof(double) * (size_t)xb_off);
    if (st == COFFE_OK) st = dev_alloc((void **)&b->d_wbuf, 6 * sizeof(double) * (size_t)b->nunit * (size_t)nsep);
    if (st == COFFE_OK) st = dev_alloc((void **)&b->d_itab, sizeof(double) * (size_t)nspec * ntab * nsep);
    /* table workgroups: consecutive units (sorted by size) of one transform length
       share a workgroup, up to the largest power of two P with threads / P >= M/8
       (each group keeps M/8 threads or more) and threads / P >= nsep (the Akima
       phase keeps one separation per thread) and 16 at most, so that the small transforms of a batch run many to a
       CU instead of one (the LDS of one workgroup holds them all) */
    int32_t *packs = NULL;
    if (st == COFFE_OK) {
        packs = (int32_t *)malloc(4 * sizeof(int32_t) * (size_t)b->nunit);
        if (!packs) st = COFFE_ENOMEM;
        const char *pe = getenv("COFFE_ENGINE_PACK");
        const int allow = !(pe && pe[0] == '0');
        b->npack = 0;
        b->lt_off = (int)CE_LDS_DATA_LEN(b->max_M);
        for (int i = 0; i < b->nunit && st == COFFE_OK;) {
            const int M = (int)(n_fft[units[i].spec] / 2);
            const int per = (M + 7) / 8;
            /* a power of two (the thread count is one): groups of exactly threads / P,
               whole wavefronts (64 threads or more) */
            int P = 1;
            while (allow && P < 16 && 2 * P * per <= b->threads && b->threads / (2 * P) >= 64 &&
                   2 * P * nsep <= b->threads)
                P *= 2;
            int c = 1;
            while (c < P && i + c < b->nunit && n_fft[units[i + c].spec] / 2 == M) ++c;
            int32_t *pk = packs + 4 * b->npack++;
            pk[0] = i;
            pk[1] = c;
            pk[2] = P;
            pk[3] = 0;
            const int need = P * CE_LDS_DATA_LEN(M);
            if (need > b->lt_off) b->lt_off = need;
            i += c;
        }
    }
    /* the ntab tables of a spectrum all read its transformed input: runs of 8 spectra
       in single-unit packs are reordered so that the packs of one spectrum get
       workgroup indices equal mod 8, i.e. one XCD (workgroups go to the 8 XCDs round
       robin), whose L2 then serves that input to all but the first table; the order
       stays largest first at the granularity of the run */
    const char *xe = getenv("COFFE_ENGINE_XCD");
    if (st == COFFE_OK && ntab > 1 && !(xe && xe[0] == '0')) {
        const int run = 8 * ntab;
        int32_t tmp[4 * 8 * CE_MAX_TABLES];
        for (int p = 0; p + run <= b->npack;) {
            int ok = 1;
            for (int q = 0; q < run && ok; ++q) {
                const int32_t *pk = packs + 4 * (p + q);
                const int32_t *head = packs + 4 * (p + (q / ntab) * ntab);
                ok = pk[1] == 1 && pk[2] == 1 && units[pk[0]].tab == q % ntab &&
                     units[pk[0]].spec == units[head[0]].spec;
            }
            if (!ok) { ++p; continue; }
            for (int q = 0; q < run; ++q) {
                const int g = q / ntab, t = q % ntab;
                memcpy(tmp + 4 * (8 * t + g), packs + 4 * (p + q), 4 * sizeof(int32_t));
            }
            memcpy(packs + 4 * p, tmp, 4 * sizeof(int32_t) * (size_t)run);
            p += run;
        }
    }
    if (st == COFFE_OK) st = dev_alloc((void **)&b->d_packs, 4 * sizeof(int32_t) * (size_t)(b->npack > 0 ? b->npack : 1));
    hipStream_t s = e->stream;
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(b->d_packs, packs, 4 * sizeof(int32_t) * (size_t)b->npack, hipMemcpyHostToDevice, s));
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(b->d_spec, b->h_spec, sizeof(ce_spec_dev) * (size_t)nspec, hipMemcpyHostToDevice, s));
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(b->d_grids, hg, sizeof(ce_grid_dev) * (size_t)b->ngrid, hipMemcpyHostToDevice, s));
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(b->d_tabs, tabs, sizeof(ce_table_desc) * (size_t)ntab, hipMemcpyHostToDevice, s));
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(b->d_units, units, sizeof(ce_unit_dev) * (size_t)b->nunit, hipMemcpyHostToDevice, s));
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(b->d_chunks, chunks, sizeof(ce_chunk_dev) * (size_t)b->nchunk, hipMemcpyHostToDevice, s));
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(b->d_sidx, sidx, sizeof(int32_t) * (size_t)b->ngrid * (size_t)nsep, hipMemcpyHostToDevice, s));
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(b->d_sep, sep, sizeof(double) * (size_t)nsep, hipMemcpyHostToDevice, s));
    if (st == COFFE_OK) st = hip_ok(hipStreamSynchronize(s));
    free(spec_grid);
    free(sidx);
    free(packs);
    free(units);
    free(chunks);
    free(hg);
    if (st != COFFE_OK) { coffe_batch_destroy(b); return st; }
    *out = b;
    return COFFE_OK;
}

static hipStream_t pick_stream(coffe_batch *b, void *stream) { return stream ? (hipStream_t)stream : b->e->stream; }

int coffe_batch_resample(coffe_batch *b, const double *d_k, const double *d_pk, void *stream)
{
    if (!b || !d_k || !d_pk) return COFFE_EINVAL;
    return ce_launch_resample(pick_stream(b, stream), b->nchunk, b->d_chunks, b->d_spec, b->d_grids, d_k, d_pk, b->d_ws,
                              b->d_ranges)
               ? COFFE_EDEVICE
               : COFFE_OK;
}

int coffe_batch_fft_prepare(coffe_batch *b, void *stream)
{
    if (!b) return COFFE_EINVAL;
    return ce_launch_prepare(pick_stream(b, stream), b->nspec, b->d_spec, b->d_grids, b->nq, b->d_ws, b->d_xf,
                             b->threads, b->max_M)
               ? COFFE_EDEVICE
               : COFFE_OK;
}

int coffe_batch_adj_c2r(coffe_batch *b, void *stream)
{
    if (!b) return COFFE_EINVAL;
    return ce_launch_adj_c2r(pick_stream(b, stream), b->nspec, b->d_spec, b->d_grids, b->d_tabs, b->ntab, b->nq,
                             b->d_ws, b->d_xb, b->threads, b->max_M, NULL, NULL, NULL, 0, 0, 0)
               ? COFFE_EDEVICE
               : COFFE_OK;
}

int coffe_batch_spline_adj(coffe_batch *b, const double *d_k, double *d_gpk, void *stream)
{
    if (!b || !d_k || !d_gpk) return COFFE_EINVAL;
    return ce_launch_spline_adj(pick_stream(b, stream), b->nchunk, b->d_chunks, b->d_spec, b->d_grids, d_k, b->d_ws,
                                d_gpk, b->d_ranges)
               ? COFFE_EDEVICE
               : COFFE_OK;
}

int coffe_batch_prepare(coffe_batch *b, const double *d_k, const double *d_pk, void *stream)
{
    if (!b || !d_k || !d_pk) return COFFE_EINVAL;
    hipStream_t s = pick_stream(b, stream);
    if (ce_launch_resample(s, b->nchunk, b->d_chunks, b->d_spec, b->d_grids, d_k, d_pk, b->d_ws, b->d_ranges))
        return COFFE_EDEVICE;
    return ce_launch_prepare(s, b->nspec, b->d_spec, b->d_grids, b->nq, b->d_ws, b->d_xf, b->threads, b->max_M)
               ? COFFE_EDEVICE
               : COFFE_OK;
}

int coffe_batch_tables_stage(coffe_batch *b, const double *d_A, const double *d_gout, double *d_itab, void *stream)
{
    if (!b || (d_gout && !d_A)) return COFFE_EINVAL;
    double *itab = d_itab ? d_itab : b->d_itab;
    return ce_launch_tables(pick_stream(b, stream), b->nunit, b->d_units, b->d_spec, b->d_grids, b->d_tabs, b->ntab,
                            b->d_xf, b->nsep, b->d_sidx, b->d_sep, itab, NULL, 0, d_gout ? d_A : NULL, d_gout,
                            b->nz, b->nl, b->d_xb, b->d_wbuf, b->threads, b->max_M, b->d_packs, b->npack, b->lt_off)
               ? COFFE_EDEVICE
               : COFFE_OK;
}

int coffe_batch_project(coffe_batch *b, const double *d_A, const double *d_itab, double *d_out, void *stream)
{
    if (!b || !d_A || !d_out) return COFFE_EINVAL;
    return ce_launch_project(pick_stream(b, stream), b->nspec, b->ntab, b->nz, b->nl, b->nsep, d_A,
                             d_itab ? d_itab : b->d_itab, d_out)
               ? COFFE_EDEVICE
               : COFFE_OK;
}

int coffe_batch_prepare_adj(coffe_batch *b, const double *d_k, const double *d_pk, double *d_gpk, void *stream)
{
    if (!b || !d_k || !d_pk || !d_gpk) return COFFE_EINVAL;
    hipStream_t s = pick_stream(b, stream);
    if (ce_launch_adj_c2r(s, b->nspec, b->d_spec, b->d_grids, b->d_tabs, b->ntab, b->nq, b->d_ws, b->d_xb,
                          b->threads, b->max_M, NULL, NULL, NULL, 0, 0, 0))
        return COFFE_EDEVICE;
    return ce_launch_spline_adj(s, b->nchunk, b->d_chunks, b->d_spec, b->d_grids, d_k, b->d_ws, d_gpk, b->d_ranges)
               ? COFFE_EDEVICE
               : COFFE_OK;
}

int coffe_batch_forward(coffe_batch *b, const double *d_k, const double *d_pk, const double *d_A, double *d_itab,
                        double *d_out, void *stream)
{
    if (!b || !d_k || !d_pk || (d_out && !d_A)) return COFFE_EINVAL;
    hipStream_t s = pick_stream(b, stream);
    double *itab = d_itab ? d_itab : b->d_itab;
    int st = coffe_batch_prepare(b, d_k, d_pk, s);
    if (st) return st;
    if (ce_launch_tables(s, b->nunit, b->d_units, b->d_spec, b->d_grids, b->d_tabs, b->ntab, b->d_xf, b->nsep,
                         b->d_sidx, b->d_sep, itab, NULL, 0, NULL, NULL, b->nz, b->nl, NULL, b->d_wbuf, b->threads,
                         b->max_M, b->d_packs, b->npack, b->lt_off))
        return COFFE_EDEVICE;
    if (d_out && ce_launch_project(s, b->nspec, b->ntab, b->nz, b->nl, b->nsep, d_A, itab, d_out)) return COFFE_EDEVICE;
    return COFFE_OK;
}

int coffe_batch_vjp(coffe_batch *b, const double *d_k, const double *d_pk, const double *d_A, const double *d_gout,
                    double *d_itab, double *d_out, double *d_gpk, void *stream)
{
    if (!b || !d_k || !d_pk || !d_A || !d_gout || !d_gpk) return COFFE_EINVAL;
    hipStream_t s = pick_stream(b, stream);
    double *itab = d_itab ? d_itab : b->d_itab;
    int st = coffe_batch_prepare(b, d_k, d_pk, s);
    if (st) return st;
    if (ce_launch_tables(s, b->nunit, b->d_units, b->d_spec, b->d_grids, b->d_tabs, b->ntab, b->d_xf, b->nsep,
                         b->d_sidx, b->d_sep, itab, NULL, 0, d_A, d_gout, b->nz, b->nl, b->d_xb, b->d_wbuf,
                         b->threads, b->max_M, b->d_packs, b->npack, b->lt_off))
        return COFFE_EDEVICE;
    /* the transposed r2c of every spectrum, with its projection (out) in the same workgroup */
    if (ce_launch_adj_c2r(s, b->nspec, b->d_spec, b->d_grids, b->d_tabs, b->ntab, b->nq, b->d_ws, b->d_xb,
                          b->threads, b->max_M, d_A, itab, d_out, b->nz, b->nl, b->nsep))
        return COFFE_EDEVICE;
    return ce_launch_spline_adj(s, b->nchunk, b->d_chunks, b->d_spec, b->d_grids, d_k, b->d_ws, d_gpk, b->d_ranges)
               ? COFFE_EDEVICE
               : COFFE_OK;
}

int coffe_batch_set_radix(coffe_batch *b, int M, int npass, const int *radix)
{
    if (!b || !radix || npass < 1 || npass > 10) return COFFE_EINVAL;
    ce_radix_plan rp;
    memset(&rp, 0, sizeof(rp));
    long long prod = 1;
    for (int p = 0; p < npass; ++p) {
        const int r = radix[p];
        if (r != 2 && r != 3 && r != 4 && r != 5 && r != 8) return COFFE_EINVAL;
        prod *= r;
        rp.radix[p] = r;
    }
    if (prod != M) return COFFE_EINVAL;
    rp.npass = npass;
    radix_code(&rp);
    hipSetDevice(b->e->device);
    int found = 0, st = COFFE_OK;
    for (int gi = 0; gi < b->ngrid && st == COFFE_OK; ++gi) {
        if (b->grids[gi]->N / 2 != M) continue;
        ce_grid_dev hg;
        grid_to_dev(b->grids[gi], &hg);
        hg.rp = rp;
        st = hip_ok(hipMemcpy(b->d_grids + gi, &hg, sizeof(hg), hipMemcpyHostToDevice));
        found = 1;
    }
    if (st != COFFE_OK) return st;
    return found ? COFFE_OK : COFFE_EDOMAIN;
}

int coffe_batch_tables(coffe_batch *b, const double *d_k, const double *d_pk, double *d_full, void *stream)
{
    if (!b || !d_k || !d_pk || !d_full) return COFFE_EINVAL;
    hipStream_t s = pick_stream(b, stream);
    int st0 = coffe_batch_prepare(b, d_k, d_pk, s);
    if (st0) return st0;
    /* the unit order is by size; the full-table rows follow the unit order, so
       launch one unit per (spec, tab) in natural order via a temporary list */
    ce_unit_dev *h = (ce_unit_dev *)malloc(sizeof(ce_unit_dev) * (size_t)b->nunit);
    ce_unit_dev *d = NULL;
    int st = h ? COFFE_OK : COFFE_ENOMEM;
    if (st == COFFE_OK) st = dev_alloc((void **)&d, sizeof(ce_unit_dev) * (size_t)b->nunit);
    if (st == COFFE_OK) {
        for (int i = 0; i < b->nunit; ++i) { h[i].spec = i / b->ntab; h[i].tab = i % b->ntab; }
        st = hip_ok(hipMemcpyAsync(d, h, sizeof(ce_unit_dev) * (size_t)b->nunit, hipMemcpyHostToDevice, s));
    }
    if (st == COFFE_OK &&
        ce_launch_tables(s, b->nunit, d, b->d_spec, b->d_grids, b->d_tabs, b->ntab, b->d_xf, b->nsep, b->d_sidx,
                         b->d_sep, NULL, d_full, b->max_N, NULL, NULL, b->nz, b->nl, NULL, b->d_wbuf, b->threads,
                         b->max_M, NULL, 0, 0))
        st = COFFE_EDEVICE;
    if (st == COFFE_OK) st = hip_ok(hipStreamSynchronize(s));
    hipFree(d);
    free(h);
    return st;
}

/* ---- drop-in: twofast_1bessel ------------------------------------------------ */

/* the scratch of at least `bytes`, for work on stream s (under e->lock): work on s
   starts after the last asynchronous kernel that read the scratch, and a reallocation
   waits for that kernel */
static int scratch_reserve(coffe_engine *e, size_t bytes, hipStream_t s)
{
    if (e->scratch_pending && hip_ok(hipStreamWaitEvent(s, e->ev_scratch, 0)) != COFFE_OK) return COFFE_EDEVICE;
    if (e->scratch_bytes >= bytes) return COFFE_OK;
    if (e->scratch_pending) {
        if (hip_ok(hipEventSynchronize(e->ev_scratch)) != COFFE_OK) return COFFE_EDEVICE;
        e->scratch_pending = 0;
    }
    hipFree(e->d_scratch);
    e->d_scratch = NULL;
    e->scratch_bytes = 0;
    if (injected_failure() || hipMalloc(&e->d_scratch, bytes) != hipSuccess) { e->d_scratch = NULL; return COFFE_ENOMEM; }
    e->scratch_bytes = bytes;
    return COFFE_OK;
}

int twofast_1bessel_status(double *output_x, double *output_y, size_t output_len, const double *input_x,
                           const double *input_y, size_t input_len, double l, double nu, double r0, double k0,
                           double kmin, double kmax)
{
    if (!output_x || !output_y || !input_x || !input_y || input_len < 5) return COFFE_EINVAL;
    if (output_len < 4 || (output_len & 1) || output_len > CE_ANY_MAX_N) return COFFE_EUNSUPPORTED;
    for (size_t i = 0; i < input_len; ++i)
        if (!isfinite(input_x[i]) || !isfinite(input_y[i])) return COFFE_ENONFINITE;
    coffe_engine *e = coffe_engine_default();
    if (!e) return COFFE_ENODEVICE;
    hipSetDevice(e->device);
    ce_table_desc td;
    td.n = nu;
    td.l = l;
    td.rpow = 0;
    td.qidx = 0;
    ce_grid *g = NULL;
    int st = grid_get(e, (int)output_len, kmin, kmax, k0, r0, 1, &td, &g);
    if (st != COFFE_OK) return st;
    const int N = (int)output_len, M = N / 2;
    const size_t n = input_len;
    /* the general path: a length-N transform of the grid, made once */
    const int general = g->rp.npass == 0;
    if (general) {
        pthread_mutex_lock(&e->lock);
        if (!g->fft) st = ce_anyfft_create(N, &g->fft);
        pthread_mutex_unlock(&e->lock);
        if (st != COFFE_OK) return st;
    }
    /* scratch layout: k, P | ws (2N) | xf (M+1 cplx; N cplx on the general path) | out (N) |
       spec, grid, tab, unit, chunks */
    const int nchunk = (int)((n - 1 + CE_CHUNK - 1) / CE_CHUNK);
    const size_t off_k = 0, off_p = n, off_ws = 2 * n, off_xf = off_ws + 2 * (size_t)N;
    const size_t off_out = off_xf + 2 * (size_t)(general ? N : M + 1), off_meta = off_out + (size_t)N;
    const size_t meta_bytes = sizeof(ce_spec_dev) + sizeof(ce_grid_dev) + sizeof(ce_table_desc) +
                              sizeof(ce_unit_dev) + sizeof(ce_chunk_dev) * (size_t)nchunk + 64;
    pthread_mutex_lock(&e->lock);
    st = scratch_reserve(e, sizeof(double) * off_meta + meta_bytes + 64, e->stream);
    if (st != COFFE_OK) { pthread_mutex_unlock(&e->lock); return st; }
    double *base = (double *)e->d_scratch;
    char *meta = (char *)(base + off_meta);
    ce_spec_dev *d_spec = (ce_spec_dev *)meta;
    ce_grid_dev *d_grid = (ce_grid_dev *)(meta + sizeof(ce_spec_dev));
    ce_table_desc *d_tab = (ce_table_desc *)((char *)d_grid + sizeof(ce_grid_dev));
    ce_unit_dev *d_unit = (ce_unit_dev *)((char *)d_tab + sizeof(ce_table_desc));
    ce_chunk_dev *d_chunks = (ce_chunk_dev *)((char *)d_unit + sizeof(ce_unit_dev));
    ce_chunk_dev *h_chunks = (ce_chunk_dev *)malloc(sizeof(ce_chunk_dev) * (size_t)nchunk);
    if (!h_chunks) { pthread_mutex_unlock(&e->lock); return COFFE_ENOMEM; }
    for (int c = 0; c < nchunk; ++c) {
        h_chunks[c].spec = 0;
        h_chunks[c].grid = 0;
        h_chunks[c].in_off = 0;
        h_chunks[c].ws_off = 0;
        h_chunks[c].n_in = (int32_t)n;
        h_chunks[c].id = c;
        h_chunks[c].a = c * CE_CHUNK;
        h_chunks[c].b = (c + 1) * CE_CHUNK < (int)n - 1 ? (c + 1) * CE_CHUNK : (int)n - 1;
    }
    ce_spec_dev hs;
    memset(&hs, 0, sizeof(hs));
    hs.in_off = 0;
    hs.n_in = (int32_t)n;
    hs.grid = 0;
    hs.xf_off = 0;
    hs.xb_off = 0;
    hs.ws_off = 0;
    ce_grid_dev hg;
    grid_to_dev(g, &hg);
    ce_unit_dev hu = {0, 0};
    hipStream_t s = e->stream;
    int threads = e->threads;
    while (threads < M / 8) threads *= 2;
    st = hip_ok(hipMemcpyAsync(base + off_k, input_x, sizeof(double) * n, hipMemcpyHostToDevice, s));
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(base + off_p, input_y, sizeof(double) * n, hipMemcpyHostToDevice, s));
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(d_spec, &hs, sizeof(hs), hipMemcpyHostToDevice, s));
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(d_grid, &hg, sizeof(hg), hipMemcpyHostToDevice, s));
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(d_tab, &g->tabs[0], sizeof(ce_table_desc), hipMemcpyHostToDevice, s));
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(d_unit, &hu, sizeof(hu), hipMemcpyHostToDevice, s));
    if (st == COFFE_OK)
        st = hip_ok(hipMemcpyAsync(d_chunks, h_chunks, sizeof(ce_chunk_dev) * (size_t)nchunk, hipMemcpyHostToDevice, s));
    if (st == COFFE_OK && ce_launch_resample(s, nchunk, d_chunks, d_spec, d_grid, base + off_k, base + off_p,
                                             base + off_ws, NULL))
        st = COFFE_EDEVICE;
    if (general) {
        /* bias * window * S -> FFT_N -> kernel multiply, Hermitian extension -> FFT_N ->
           real part times the output prefactor (twofast.c:362-483) */
        const int qi = g->tabs[0].qidx;
        if (st == COFFE_OK && ce_launch_tf_in(s, base + off_ws, g->d_bw[qi], base + off_xf, N)) st = COFFE_EDEVICE;
        if (st == COFFE_OK) st = ce_anyfft_forward(g->fft, base + off_xf, s);
        if (st == COFFE_OK && ce_launch_tf_mid(s, base + off_xf, g->d_U[0], N)) st = COFFE_EDEVICE;
        if (st == COFFE_OK) st = ce_anyfft_forward(g->fft, base + off_xf, s);
        if (st == COFFE_OK && ce_launch_tf_out(s, base + off_xf, g->d_os[0], base + off_out, N)) st = COFFE_EDEVICE;
    } else {
        if (st == COFFE_OK && ce_launch_prepare(s, 1, d_spec, d_grid, 1, base + off_ws, base + off_xf, threads, M))
            st = COFFE_EDEVICE;
        if (st == COFFE_OK && ce_launch_tables(s, 1, d_unit, d_spec, d_grid, d_tab, 1, base + off_xf, 0, NULL, NULL,
                                               NULL, base + off_out, N, NULL, NULL, 1, 1, NULL, NULL, threads, M,
                                               NULL, 0, 0))
            st = COFFE_EDEVICE;
    }
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(output_y, base + off_out, sizeof(double) * (size_t)N, hipMemcpyDeviceToHost, s));
    if (st == COFFE_OK) st = hip_ok(hipStreamSynchronize(s));
    pthread_mutex_unlock(&e->lock);
    free(h_chunks);
    if (st != COFFE_OK) return st;
    for (int i = 0; i < N; ++i) output_x[i] = r0 * g->h_p[i];
    return COFFE_OK;
}

int coffe_fft_any(const double *in, double *out, int64_t L)
{
    if (!in || !out || L < 1) return COFFE_EINVAL;
    coffe_engine *e = coffe_engine_default();
    if (!e) return COFFE_ENODEVICE;
    hipSetDevice(e->device);
    ce_anyfft *p = NULL;
    int st = ce_anyfft_create(L, &p);
    if (st != COFFE_OK) return st;
    void *d = NULL;
    st = dev_alloc(&d, sizeof(double) * 2 * (size_t)L);
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(d, in, sizeof(double) * 2 * (size_t)L, hipMemcpyHostToDevice, e->stream));
    if (st == COFFE_OK) st = ce_anyfft_forward(p, d, e->stream);
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(out, d, sizeof(double) * 2 * (size_t)L, hipMemcpyDeviceToHost, e->stream));
    if (st == COFFE_OK) st = hip_ok(hipStreamSynchronize(e->stream));
    hipFree(d);
    ce_anyfft_destroy(p);
    return st;
}

void twofast_1bessel(double *output_x, double *output_y, size_t output_len, double *input_x, double *input_y,
                     size_t input_len, double l, double nu, double r0, double k0, double kmin, double kmax,
                     unsigned flag)
{
    (void)flag;
    ce_last_error = twofast_1bessel_status(output_x, output_y, output_len, input_x, input_y, input_len, l, nu, r0,
                                           k0, kmin, kmax);
}

/* ---- mu-quadrature weights ---------------------------------------------------- */

/* Gauss-Legendre nodes/weights on [-1, 1] by Newton iteration on the Legendre
   recurrence in long double */
static void gauss_legendre(int n, double *x, double *w)
{
    for (int i = 0; i < (n + 1) / 2; ++i) {
        long double z = cosl(3.14159265358979323846264338327950288L * ((long double)i + 0.75L) / ((long double)n + 0.5L));
        long double pp = 1.0L;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1.0L, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const long double p2 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p0) / (long double)k;
                p0 = p1;
                p1 = p2;
            }
            pp = (long double)n * (z * p1 - p0) / (z * z - 1.0L);
            const long double dz = p1 / pp;
            z -= dz;
            if (fabsl(dz) < 1e-19L) break;
        }
        long double p0 = 1.0L, p1 = z;
        for (int k = 2; k <= n; ++k) {
            const long double p2 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p0) / (long double)k;
            p0 = p1;
            p1 = p2;
        }
        pp = (long double)n * (z * p1 - p0) / (z * z - 1.0L);
        const long double ww = 2.0L / ((1.0L - z * z) * pp * pp);
        x[i] = -(double)z;
        x[n - 1 - i] = (double)z;
        w[i] = (double)ww;
        w[n - 1 - i] = (double)ww;
    }
    if (n & 1) x[n / 2] = 0.0;
}

/* Gauss-Legendre rules, nodes then weights (2 n doubles), computed once per order for
   the first 64 orders a process asks for (kept until exit); later orders are computed
   per call into a buffer the caller frees (*owned = 1) */
static const double *gl_nodes(int n, int *owned)
{
    static pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
    static struct { int n; double *xw; } cache[64];
    static int ncache = 0;
    const double *out = NULL;
    *owned = 0;
    pthread_mutex_lock(&mu);
    for (int i = 0; i < ncache && !out; ++i)
        if (cache[i].n == n) out = cache[i].xw;
    if (!out) {
        double *xw = (double *)malloc(sizeof(double) * 2 * (size_t)n);
        if (xw) {
            gauss_legendre(n, xw, xw + n);
            if (ncache < 64) { cache[ncache].n = n; cache[ncache].xw = xw; ++ncache; }
            else *owned = 1;
            out = xw;
        }
    }
    pthread_mutex_unlock(&mu);
    return out;
}

static int std_slot(int n, int l)
{
    static const int tabs[8][2] = {{0, 0}, {0, 2}, {0, 4}, {1, 1}, {1, 3}, {2, 0}, {2, 2}, {3, 1}};
    for (int s = 0; s < 8; ++s)
        if (tabs[s][0] == n && tabs[s][1] == l) return s;
    return -1;
}

unsigned coffe_terms_of(int den, int rsd, int d1)
{
    unsigned t = 0;
    if (den) t |= COFFE_TERM_DEN_DEN;
    if (rsd) t |= COFFE_TERM_RSD_RSD;
    if (d1) t |= COFFE_TERM_D1_D1;
    if (den && rsd) t |= COFFE_TERM_DEN_RSD;
    if (den && d1) t |= COFFE_TERM_DEN_D1;
    if (rsd && d1) t |= COFFE_TERM_RSD_D1;
    return t;
}

int coffe_mu_weights(coffe_engine *e, int nbg, const double *bg_z, const double *bg_chi, const double *bg_D1,
                     const double *bg_f, const double *bg_cH, const double *bg_G1, const double *bg_G2, int nb,
                     const double *bias_z, const double *bias1, const double *bias2, int den, int rsd, int d1,
                     int flatsky, int nz, const double *zmean, int nl, const int *ls, int ntab, const int *tab_n,
                     const int *tab_l, int nsep, const double *sep, int nmu, double *d_A, void *stream)
{
    return coffe_mu_weights_terms(e, nbg, bg_z, bg_chi, bg_D1, bg_f, bg_cH, bg_G1, bg_G2, nb, bias_z, bias1, bias2,
                                  coffe_terms_of(den, rsd, d1), flatsky, nz, zmean, nl, ls, ntab, tab_n, tab_l, nsep,
                                  sep, nmu, d_A, stream);
}

int coffe_mu_weights_terms(coffe_engine *e, int nbg, const double *bg_z, const double *bg_chi, const double *bg_D1,
                           const double *bg_f, const double *bg_cH, const double *bg_G1, const double *bg_G2, int nb,
                           const double *bias_z, const double *bias1, const double *bias2, unsigned terms,
                           int flatsky, int nz, const double *zmean, int nl, const int *ls, int ntab,
                           const int *tab_n, const int *tab_l, int nsep, const double *sep, int nmu, double *d_A,
                           void *stream)
{
    if (terms > 63) return COFFE_EINVAL;   /* the d2 / g terms need coffe_mu_weights_ex */
    coffe_mu_inputs in;
    memset(&in, 0, sizeof(in));
    in.nbg = nbg; in.bg_z = bg_z; in.bg_chi = bg_chi; in.bg_D1 = bg_D1; in.bg_f = bg_f; in.bg_cH = bg_cH;
    in.bg_G1 = bg_G1; in.bg_G2 = bg_G2;
    in.nb = nb; in.bias_z = bias_z; in.bias1 = bias1; in.bias2 = bias2;
    return coffe_mu_weights_ex(e, &in, terms, 0, flatsky, 0, nz, zmean, nl, ls, ntab, tab_n, tab_l, nsep, sep, nmu,
                               d_A, NULL, stream);
}

/* node slopes of the natural cubic spline through (x_i, y_i), i < n (the derivative
   tables of GSL's bicubic): second derivatives M from the tridiagonal system with M_0 =
   M_{n-1} = 0 (Thomas elimination), then y'_i from the interval formulas; y and out
   are strided, work holds 2 n doubles */
static void spline_node_slopes(int n, const double *x, const double *y, int ys, double *out, int os, double *work)
{
    double *M = work, *cp = work + n;
    M[0] = 0.0;
    cp[0] = 0.0;
    /* forward elimination of h_{i-1} M_{i-1} + 2 (h_{i-1} + h_i) M_i + h_i M_{i+1} = r_i */
    for (int i = 1; i < n - 1; ++i) {
        const double h0 = x[i] - x[i - 1], h1 = x[i + 1] - x[i];
        const double r = 6.0 * ((y[(i + 1) * ys] - y[i * ys]) / h1 - (y[i * ys] - y[(i - 1) * ys]) / h0);
        const double den = 2.0 * (h0 + h1) - h0 * cp[i - 1];
        cp[i] = h1 / den;
        M[i] = (r - h0 * M[i - 1]) / den;
    }
    M[n - 1] = 0.0;
    for (int i = n - 2; i >= 1; --i) M[i] -= cp[i] * M[i + 1];
    for (int i = 0; i < n - 1; ++i) {
        const double h = x[i + 1] - x[i];
        out[i * os] = (y[(i + 1) * ys] - y[i * ys]) / h - h * (2.0 * M[i] + M[i + 1]) / 6.0;
    }
    const double h = x[n - 1] - x[n - 2];
    out[(n - 1) * os] = (y[(n - 1) * ys] - y[(n - 2) * ys]) / h + h * (M[n - 2] + 2.0 * M[n - 1]) / 6.0;
}

/* the mu rule: rule_mu / rule_w (nmu nodes) or, when NULL, the nmu-node Gauss-Legendre rule */
static int mu_weights_impl(coffe_engine *e, const coffe_mu_inputs *in, unsigned terms, unsigned sources, int flatsky,
                           int paired, int nz, const double *zmean, int nl, const int *ls, int ntab, const int *tab_n,
                           const int *tab_l, int nsep, const double *sep, int nmu, const double *rule_mu,
                           const double *rule_w, double *d_A, double *d_ren, const coffe_mu_inputs *tan,
                           double *d_At, void *stream)
{
    if (!e || !in || nz <= 0 || nl <= 0 || nl > 4 || ntab <= 0 || ntab > CE_MAX_TABLES + 1 || nsep <= 0 ||
        nmu < (rule_mu ? 1 : 2) || nmu > (1 << 20) || (!rule_mu) != (!rule_w) ||
        terms == 0 || terms > 2047 || !d_A || !zmean || !ls || !tab_n || !tab_l || !sep || (paired && nz != nsep))
        return COFFE_EINVAL;
    /* the mutual-only d2 / g terms qualify COFFE_TERM_DG_DG */
    if ((terms & COFFE_TERM_DG_MUTUAL) && !(terms & COFFE_TERM_DG_DG)) return COFFE_EINVAL;
    const int nbg = in->nbg, nb = in->nb;
    if (nbg < 5 || nb < 5 || !in->bg_z || !in->bg_chi || !in->bg_D1 || !in->bg_f || !in->bg_cH || !in->bg_G1 ||
        !in->bg_G2 || !in->bias_z || !in->bias1 || !in->bias2)
        return COFFE_EINVAL;
    const unsigned dgterms = COFFE_TERM_DG_DG | COFFE_TERM_DEN_DG | COFFE_TERM_RSD_DG | COFFE_TERM_D1_DG;
    const int dg = (terms & dgterms) != 0;
    if (dg != (sources != 0) || sources > 15) return COFFE_EINVAL;
    /* the tangent instance: density / RSD / Doppler-1 terms, the six background tangents */
    if (tan && (dg || !d_At || !tan->bg_chi || !tan->bg_D1 || !tan->bg_f || !tan->bg_cH || !tan->bg_G1 ||
                !tan->bg_G2 || ntab > CE_MAX_TABLES))
        return dg ? COFFE_EUNSUPPORTED : COFFE_EINVAL;
    const int nmb = dg ? in->nmb : 0;
    if (dg && (nmb < 5 || !in->bg_a || !in->mb_z || !in->s1 || !in->s2 || !in->fevo1 || !in->fevo2 ||
               !(in->H0h > 0)))
        return COFFE_EINVAL;
    const int ren = (terms & COFFE_TERM_DG_DG) != 0;
    const int nren = ren ? in->nren : 0;
    if (ren && (nren < 5 || !in->ren_chi || !in->ren || !d_ren)) return COFFE_EINVAL;
    int slot[CE_MAX_TABLES + 1], have[9] = {0};
    for (int t = 0; t < ntab; ++t) {
        slot[t] = (tab_n[t] == 4 && tab_l[t] == 0) ? 8 : std_slot(tab_n[t], tab_l[t]);
        if (slot[t] < 0) return COFFE_EINVAL;
        have[slot[t]] = 1;
    }
    /* the tables each contribution needs (functions.c:129-1428) */
    if ((terms & COFFE_TERM_DEN_DEN) && !have[0]) return COFFE_EINVAL;
    if ((terms & COFFE_TERM_RSD_RSD) && !(have[0] && have[1] && have[2])) return COFFE_EINVAL;
    if ((terms & COFFE_TERM_D1_D1) && !(have[5] && have[6])) return COFFE_EINVAL;
    if ((terms & COFFE_TERM_DEN_RSD) && !(have[0] && have[1])) return COFFE_EINVAL;
    if ((terms & COFFE_TERM_DEN_D1) && !have[3]) return COFFE_EINVAL;
    if ((terms & COFFE_TERM_RSD_D1) && !(have[3] && have[4])) return COFFE_EINVAL;
    if ((terms & COFFE_TERM_DG_DG) && !have[8]) return COFFE_EINVAL;
    if ((terms & COFFE_TERM_DEN_DG) && !have[5]) return COFFE_EINVAL;
    if ((terms & COFFE_TERM_RSD_DG) && !(have[5] && have[6])) return COFFE_EINVAL;
    if ((terms & COFFE_TERM_D1_DG) && !have[7]) return COFFE_EINVAL;
    for (int i = 0; i < nl; ++i)
        if (ls[i] < 0) return COFFE_EINVAL;
    /* chi(z_mean) -/+ sep/2 must stay inside the chi table (and the renormalisation grid)
       at every |mu| <= 1 and z_mean inside the bias tables (the reference stops on a GSL
       domain error) */
    for (int iz = 0; iz < nz; ++iz)
        if (zmean[iz] < in->bg_z[0] || zmean[iz] > in->bg_z[nbg - 1] || zmean[iz] < in->bias_z[0] ||
            zmean[iz] > in->bias_z[nb - 1])
            return COFFE_EDOMAIN;
    for (int r = 0; r < nsep; ++r)
        if (!(sep[r] > 0.0)) return COFFE_EDOMAIN;
    double *chim = (double *)malloc(sizeof(double) * (size_t)nz);
    if (!chim) return COFFE_ENOMEM;
    if (coffe_akima_eval((size_t)nbg, in->bg_z, in->bg_chi, (size_t)nz, zmean, chim) != COFFE_OK) {
        free(chim);
        return COFFE_EINVAL;
    }
    double smax = 0.0;
    for (int r = 0; r < nsep; ++r) smax = sep[r] > smax ? sep[r] : smax;
    const double lo = ren ? fmax(in->bg_chi[0], in->ren_chi[0]) : in->bg_chi[0];
    const double hi = ren ? fmin(in->bg_chi[nbg - 1], in->ren_chi[nren - 1]) : in->bg_chi[nbg - 1];
    for (int iz = 0; iz < nz; ++iz) {
        const double s = paired ? sep[iz] : smax;
        if (!(chim[iz] - s / 2 >= lo && chim[iz] + s / 2 <= hi)) { free(chim); return COFFE_EDOMAIN; }
    }
    free(chim);
    hipSetDevice(e->device);
    hipStream_t s = stream ? (hipStream_t)stream : e->stream;
    const size_t nnode = (size_t)nren * (size_t)nren;
    /* host arrays staged in one pinned buffer: 8 background tables, 3 bias tables, the 5
       magnification / evolution tables, zmean, sep, the rule (2 nmu), the renormalisation
       chi grid and its values with their three derivative tables */
    const size_t ntan = tan ? 6 * (size_t)nbg : 0;
    const size_t nstage = 8 * (size_t)nbg + 3 * (size_t)nb + 5 * (size_t)nmb + (size_t)nz + (size_t)nsep +
                          2 * (size_t)nmu + (size_t)nren + 4 * nnode + ntan;
    const size_t ncoef = 3 * (7 * (size_t)nbg + 2 * (size_t)nb + 4 * (size_t)nmb) + 3 * ntan;
    const size_t total = sizeof(double) * (nstage + ncoef);
    int owned = 0;
    const double *h_mu = NULL;
    if (rule_mu) {
        double *r = (double *)malloc(sizeof(double) * 2 * (size_t)nmu);
        if (r) {
            memcpy(r, rule_mu, sizeof(double) * (size_t)nmu);
            memcpy(r + nmu, rule_w, sizeof(double) * (size_t)nmu);
            owned = 1;
        }
        h_mu = r;
    } else {
        h_mu = gl_nodes(nmu, &owned);
    }
    if (!h_mu) return COFFE_ENOMEM;
    /* the inputs are staged in the engine's scratch (the drop-in calls share it under the
       lock).  The call is asynchronous on the stream: the next user of the scratch waits
       for this kernel (ev_scratch) and the next writer of this pinned staging buffer for
       this copy (ev_stage) */
    pthread_mutex_lock(&e->lock);
    int st = scratch_reserve(e, total, s);
    if (st != COFFE_OK) { pthread_mutex_unlock(&e->lock); if (owned) free((void *)h_mu); return st; }
    const int sb = e->stage_next;
    e->stage_next ^= 1;
    if (e->stage_pending[sb]) {
        st = hip_ok(hipEventSynchronize(e->ev_stage[sb]));
        e->stage_pending[sb] = 0;
    }
    if (st == COFFE_OK && e->pinned_bytes[sb] < sizeof(double) * nstage) {
        hipHostFree(e->h_pinned[sb]);
        e->h_pinned[sb] = NULL;
        e->pinned_bytes[sb] = 0;
        if (hipHostMalloc(&e->h_pinned[sb], sizeof(double) * nstage, hipHostMallocDefault) == hipSuccess)
            e->pinned_bytes[sb] = sizeof(double) * nstage;
        else
            st = COFFE_ENOMEM;
    }
    ce_mu_dev d;
    memset(&d, 0, sizeof(d));
    double *dp = (double *)e->d_scratch, *hp = (double *)e->h_pinned[sb];
    /* stage n values of src (or zeros when src is NULL), return the device copy */
#define CE_STAGE(src, n)                                                                 \
    ({                                                                                   \
        const size_t nn_ = (size_t)(n);                                                  \
        const double *src_ = (src);                                                      \
        if (src_) memcpy(hp, src_, sizeof(double) * nn_);                                \
        else memset(hp, 0, sizeof(double) * nn_);                                        \
        const double *r_ = dp;                                                           \
        dp += nn_;                                                                       \
        hp += nn_;                                                                       \
        r_;                                                                              \
    })
    double *d_z = NULL, *d_sep = NULL, *d_mu = NULL;
    ce_mu_tan dt;
    memset(&dt, 0, sizeof(dt));
    if (st == COFFE_OK) {
        d.nbg = nbg;
        d.bg_z = CE_STAGE(in->bg_z, nbg);
        d.bg_chi = CE_STAGE(in->bg_chi, nbg);
        d.bg_D1 = CE_STAGE(in->bg_D1, nbg);
        d.bg_f = CE_STAGE(in->bg_f, nbg);
        d.bg_cH = CE_STAGE(in->bg_cH, nbg);
        d.bg_G1 = CE_STAGE(in->bg_G1, nbg);
        d.bg_G2 = CE_STAGE(in->bg_G2, nbg);
        d.bg_a = CE_STAGE(dg ? in->bg_a : NULL, nbg);
        d.nb = nb;
        d.bias_z = CE_STAGE(in->bias_z, nb);
        d.bias1 = CE_STAGE(in->bias1, nb);
        d.bias2 = CE_STAGE(in->bias2, nb);
        d.nmb = nmb;
        if (dg) {
            d.mb_z = CE_STAGE(in->mb_z, nmb);
            d.s1 = CE_STAGE(in->s1, nmb);
            d.s2 = CE_STAGE(in->s2, nmb);
            d.fevo1 = CE_STAGE(in->fevo1, nmb);
            d.fevo2 = CE_STAGE(in->fevo2, nmb);
        }
        d.H0h = in->H0h;
        d.Om = in->Omega_m;
        if (tan) {
            dt.chi = CE_STAGE(tan->bg_chi, nbg);
            dt.D1 = CE_STAGE(tan->bg_D1, nbg);
            dt.f = CE_STAGE(tan->bg_f, nbg);
            dt.cH = CE_STAGE(tan->bg_cH, nbg);
            dt.G1 = CE_STAGE(tan->bg_G1, nbg);
            dt.G2 = CE_STAGE(tan->bg_G2, nbg);
        }
        d_z = (double *)CE_STAGE(zmean, nz);
        d_sep = (double *)CE_STAGE(sep, nsep);
        d_mu = (double *)CE_STAGE(h_mu, 2 * (size_t)nmu);
        if (ren) {
            d.nren = nren;
            d.ren_chi = CE_STAGE(in->ren_chi, nren);
            /* the values and GSL bicubic's derivative tables: natural-spline slopes along
               x (each row), along y (each column), and along y of the x slopes */
            double *hv = hp, *hx = hp + nnode, *hy = hp + 2 * nnode, *hxy = hp + 3 * nnode;
            double *work = (double *)malloc(sizeof(double) * 2 * (size_t)nren);
            if (!work) st = COFFE_ENOMEM;
            if (st == COFFE_OK) {
                memcpy(hv, in->ren, sizeof(double) * nnode);
                for (int j = 0; j < nren; ++j) spline_node_slopes(nren, in->ren_chi, hv + (size_t)j * nren, 1, hx + (size_t)j * nren, 1, work);
                for (int i = 0; i < nren; ++i) spline_node_slopes(nren, in->ren_chi, hv + i, nren, hy + i, nren, work);
                for (int i = 0; i < nren; ++i) spline_node_slopes(nren, in->ren_chi, hx + i, nren, hxy + i, nren, work);
            }
            free(work);
            d.ren_v = dp;
            d.ren_x = dp + nnode;
            d.ren_y = dp + 2 * nnode;
            d.ren_xy = dp + 3 * nnode;
            dp += 4 * nnode;
            hp += 4 * nnode;
        }
    }
#undef CE_STAGE
    if (st == COFFE_OK) {
        st = hip_ok(hipMemcpyAsync(e->d_scratch, e->h_pinned[sb], sizeof(double) * nstage, hipMemcpyHostToDevice, s));
        if (st == COFFE_OK && hip_ok(hipEventRecord(e->ev_stage[sb], s)) == COFFE_OK) e->stage_pending[sb] = 1;
    }
    double *d_coef = (double *)e->d_scratch + nstage;
    if (st == COFFE_OK && tan) {
        if (ce_launch_mu_weights_jvp(s, &d, &dt, terms, flatsky, paired, nz, d_z, nl, ls, ntab, slot, nsep, d_sep,
                                     nmu, d_mu, d_mu + nmu, d_A, d_At, d_coef))
            st = COFFE_EDEVICE;
    } else if (st == COFFE_OK &&
               ce_launch_mu_weights(s, &d, terms, sources, flatsky, paired, nz, d_z, nl, ls, ntab, slot, nsep, d_sep,
                                    nmu, d_mu, d_mu + nmu, d_A, ren ? d_ren : NULL, d_coef)) {
        st = COFFE_EDEVICE;
    }
    if (st == COFFE_OK) {
        st = hip_ok(hipEventRecord(e->ev_scratch, s));
        e->scratch_pending = (st == COFFE_OK);
    }
    if (st != COFFE_OK) hipStreamSynchronize(s);   /* nothing of a failed call stays in flight */
    pthread_mutex_unlock(&e->lock);
    if (owned) free((void *)h_mu);
    return st;
}

int coffe_mu_weights_ex(coffe_engine *e, const coffe_mu_inputs *in, unsigned terms, unsigned sources, int flatsky,
                        int paired, int nz, const double *zmean, int nl, const int *ls, int ntab, const int *tab_n,
                        const int *tab_l, int nsep, const double *sep, int nmu, double *d_A, double *d_ren,
                        void *stream)
{
    return mu_weights_impl(e, in, terms, sources, flatsky, paired, nz, zmean, nl, ls, ntab, tab_n, tab_l, nsep, sep,
                           nmu, NULL, NULL, d_A, d_ren, NULL, NULL, stream);
}

int coffe_mu_weights_jvp(coffe_engine *e, const coffe_mu_inputs *in, const coffe_mu_inputs *tangent, unsigned terms,
                         int flatsky, int paired, int nz, const double *zmean, int nl, const int *ls, int ntab,
                         const int *tab_n, const int *tab_l, int nsep, const double *sep, int nmu, double *d_A,
                         double *d_At, void *stream)
{
    if (!tangent || !d_At) return COFFE_EINVAL;
    return mu_weights_impl(e, in, terms, 0, flatsky, paired, nz, zmean, nl, ls, ntab, tab_n, tab_l, nsep, sep, nmu,
                           NULL, NULL, d_A, NULL, tangent, d_At, stream);
}

int coffe_mu_weights_rule(coffe_engine *e, const coffe_mu_inputs *in, unsigned terms, unsigned sources, int flatsky,
                          int paired, int nz, const double *zmean, int nl, const int *ls, int ntab, const int *tab_n,
                          const int *tab_l, int nsep, const double *sep, int nmu, const double *mu,
                          const double *wmu, double *d_A, double *d_ren, void *stream)
{
    if (!mu || !wmu) return COFFE_EINVAL;
    for (int q = 0; q < nmu; ++q)
        if (!(mu[q] >= -1.0 && mu[q] <= 1.0) || !isfinite(wmu[q])) return COFFE_EDOMAIN;
    return mu_weights_impl(e, in, terms, sources, flatsky, paired, nz, zmean, nl, ls, ntab, tab_n, tab_l, nsep, sep,
                           nmu, mu, wmu, d_A, d_ren, NULL, NULL, stream);
}

/* ---- small-separation nodes --------------------------------------------------- */

static int check_small_r_status(const int *h_stat, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        if (h_stat[i] == 3) return COFFE_ENONFINITE;
    for (size_t i = 0; i < n; ++i)
        if (h_stat[i] != 0) return COFFE_ENOCONV;
    return COFFE_OK;
}

int coffe_batch_small_r(coffe_batch *b, const double *d_k, const double *d_pk, double *d_out, void *stream)
{
    if (!b || !d_k || !d_pk || !d_out) return COFFE_EINVAL;
    hipSetDevice(b->e->device);
    const size_t n = (size_t)b->nspec * (size_t)b->ntab * (1 + CE_MIN_SEP_LEN);
    if (!b->d_srstat && dev_alloc((void **)&b->d_srstat, sizeof(int) * n) != COFFE_OK) return COFFE_ENOMEM;
    hipStream_t s = stream ? (hipStream_t)stream : b->e->stream;
    if (ce_launch_small_r(s, b->nspec, b->ntab, b->d_spec, b->d_grids, b->d_tabs, d_k, d_pk, 1e-5, d_out,
                          b->d_srstat))
        return COFFE_EDEVICE;
    int *h = (int *)malloc(sizeof(int) * n);
    if (!h) return COFFE_ENOMEM;
    int st = hip_ok(hipMemcpyAsync(h, b->d_srstat, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    if (st == COFFE_OK) st = hip_ok(hipStreamSynchronize(s));
    if (st == COFFE_OK) st = check_small_r_status(h, n);
    free(h);
    return st;
}

/* the reference's coffe_interpolation (common.h:228-232) and the head of GSL's
   gsl_spline it points to (interp, x, y, size: gsl_spline.h, GSL 2.7) */
typedef struct {
    void *interp;
    double *x;
    double *y;
    size_t size;
} ce_gsl_spline_view;
typedef struct {
    ce_gsl_spline_view *spline;
    void *accel;
} ce_interpolation_view;

int coffe_integrals_renormalizable(double **output_x, double **output_y, const size_t output_len,
                                   size_t *real_output_len, const struct coffe_interpolation *spectrum, const int n,
                                   const int l, const int state_n, const int state_l, const double x_min,
                                   const double x_max)
{
    if (!output_x || !output_y || !real_output_len || !spectrum) return COFFE_EINVAL;
    const ce_interpolation_view *sp = (const ce_interpolation_view *)spectrum;
    if (!sp->spline || !sp->spline->x || !sp->spline->y) return COFFE_EINVAL;
    const size_t len = output_len + 1 + CE_MIN_SEP_LEN;
    double *x = (double *)malloc(sizeof(double) * len);
    double *y = (double *)malloc(sizeof(double) * len);
    if (!x || !y) {
        free(x);
        free(y);
        return COFFE_ENOMEM;
    }
    const int st = coffe_integral_table_state(n, l, state_n, state_l, output_len, sp->spline->x, sp->spline->y,
                                              sp->spline->size, x_min, x_max, x, y);
    if (st != COFFE_OK) {
        free(x);
        free(y);
        return st;
    }
    *output_x = x;
    *output_y = y;
    *real_output_len = len;
    return COFFE_OK;
}

int coffe_divergent_tables_device(coffe_engine *e, const double *d_k, const double *d_pk, size_t nin, double kmin,
                                  double kmax, const double *d_sep, size_t nsep, const double *d_chi, size_t nchi,
                                  double *d_i40, double *d_r0, double *d_ren, int *d_status, int *d_intervals,
                                  void *stream)
{
    if (!e || !d_k || !d_pk || nin < 5 || !d_status || (nsep && (!d_sep || !d_i40 || !d_r0)) ||
        (nchi && (!d_chi || !d_ren)))
        return COFFE_EINVAL;
    if (!(kmin > 0) || !(kmin < kmax) || !isfinite(kmax)) return COFFE_EDOMAIN;
    if (nsep > (1u << 20) || nchi > 4096 || nin > ((size_t)1 << 28)) return COFFE_EINVAL;
    if (2 * nsep + nchi * nchi == 0) return COFFE_OK;
    hipSetDevice(e->device);
    hipStream_t s = stream ? (hipStream_t)stream : e->stream;
    if (ce_launch_divergent(s, d_k, d_pk, (int)nin, kmin, kmax, d_sep, (int)nsep, d_chi, (int)nchi, 1e-5, d_i40,
                            d_r0, d_ren, d_status, d_intervals))
        return COFFE_EDEVICE;
    return COFFE_OK;
}

int coffe_divergent_tables(const double *k, const double *pk, size_t nin, double kmin, double kmax,
                           const double *sep, size_t nsep, const double *chi, size_t nchi, double *out_i40,
                           double *out_r0, double *out_ren, int *out_status, int *out_intervals)
{
    if (!k || !pk || nin < 5 || (nsep && (!sep || !out_i40 || !out_r0)) || (nchi && (!chi || !out_ren)))
        return COFFE_EINVAL;
    if (!(kmin > 0) || !(kmin >= k[0]) || !(kmax <= k[nin - 1]) || !(kmin < kmax)) return COFFE_EDOMAIN;
    if (nsep > (1u << 20) || nchi > 4096 || nin > ((size_t)1 << 28)) return COFFE_EINVAL;
    for (size_t i = 0; i < nsep; ++i)
        if (!(sep[i] >= 0) || !isfinite(sep[i])) return COFFE_EDOMAIN;
    for (size_t i = 0; i < nchi; ++i)
        if (!(chi[i] >= 0) || !isfinite(chi[i])) return COFFE_EDOMAIN;
    const size_t units = 2 * nsep + nchi * nchi;
    if (units == 0) return COFFE_OK;
    coffe_engine *e = coffe_engine_default();
    if (!e) return COFFE_ENODEVICE;
    hipSetDevice(e->device);
    /* scratch: k, P | sep | chi | i40 | r0 | ren | status | interval counts */
    const size_t o_sep = 2 * nin, o_chi = o_sep + nsep, o_i40 = o_chi + nchi, o_r0 = o_i40 + nsep,
                 o_ren = o_r0 + nsep, o_stat = o_ren + nchi * nchi;
    int *h_stat = (int *)malloc(sizeof(int) * units);
    const int want_n = out_intervals != NULL;
    if (!h_stat) return COFFE_ENOMEM;
    pthread_mutex_lock(&e->lock);
    int st = scratch_reserve(e, sizeof(double) * o_stat + 2 * sizeof(int) * units + 64, e->stream);
    if (st != COFFE_OK) { pthread_mutex_unlock(&e->lock); free(h_stat); return st; }
    double *base = (double *)e->d_scratch;
    int *d_stat = (int *)(base + o_stat);
    int *d_nint = d_stat + units;
    hipStream_t s = e->stream;
    st = hip_ok(hipMemcpyAsync(base, k, sizeof(double) * nin, hipMemcpyHostToDevice, s));
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(base + nin, pk, sizeof(double) * nin, hipMemcpyHostToDevice, s));
    if (st == COFFE_OK && nsep)
        st = hip_ok(hipMemcpyAsync(base + o_sep, sep, sizeof(double) * nsep, hipMemcpyHostToDevice, s));
    if (st == COFFE_OK && nchi)
        st = hip_ok(hipMemcpyAsync(base + o_chi, chi, sizeof(double) * nchi, hipMemcpyHostToDevice, s));
    if (st == COFFE_OK)
        st = coffe_divergent_tables_device(e, base, base + nin, nin, kmin, kmax, base + o_sep, nsep, base + o_chi, nchi,
                                           base + o_i40, base + o_r0, base + o_ren, d_stat, want_n ? d_nint : NULL, s);
    if (st == COFFE_OK && nsep) {
        st = hip_ok(hipMemcpyAsync(out_i40, base + o_i40, sizeof(double) * nsep, hipMemcpyDeviceToHost, s));
        if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(out_r0, base + o_r0, sizeof(double) * nsep, hipMemcpyDeviceToHost, s));
    }
    if (st == COFFE_OK && nchi)
        st = hip_ok(hipMemcpyAsync(out_ren, base + o_ren, sizeof(double) * nchi * nchi, hipMemcpyDeviceToHost, s));
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(h_stat, d_stat, sizeof(int) * units, hipMemcpyDeviceToHost, s));
    if (st == COFFE_OK && want_n)
        st = hip_ok(hipMemcpyAsync(out_intervals, d_nint, sizeof(int) * units, hipMemcpyDeviceToHost, s));
    if (st == COFFE_OK) st = hip_ok(hipStreamSynchronize(s));
    if (st != COFFE_OK) hipStreamSynchronize(s);
    pthread_mutex_unlock(&e->lock);
    if (st == COFFE_OK) {
        if (out_status) memcpy(out_status, h_stat, sizeof(int) * units);
        /* the reference runs these integrals with GSL's error handler off
           (integrals.c:593) and keeps roundoff-limited results; only a non-finite
           value or this engine's interval cap (1024, the reference's 50000) fails */
        for (size_t i = 0; i < units && st == COFFE_OK; ++i)
            if (h_stat[i] == 4) st = COFFE_ENONFINITE;
        for (size_t i = 0; i < units && st == COFFE_OK; ++i)
            if (h_stat[i] == 1) st = COFFE_ENOCONV;
    }
    free(h_stat);
    return st;
}

/* the table of (n, l), both integers or both half-integers (integrals.c:186-207) */
static int integral_table_any(double n, double l, size_t N, const double *k, const double *pk, size_t nin,
                              double kmin, double kmax, double *out_x, double *out_y)
{
    const size_t nodes = 1 + CE_MIN_SEP_LEN;
    if (!(kmin >= k[0]) || !(kmax <= k[nin - 1]) || !(kmin < kmax)) return COFFE_EDOMAIN;
    /* FFT nodes (integrals.c:496-522) */
    int st = twofast_1bessel_status(out_x + nodes, out_y + nodes, N, k, pk, nin, l, n, 1.0, kmin, kmin, kmax);
    if (st != COFFE_OK) return st;
    if (n >= l)
        for (size_t i = 0; i < N; ++i) out_y[nodes + i] *= pow(out_x[nodes + i], n - l);
    /* small-r nodes (integrals.c:524-557) */
    coffe_engine *e = coffe_engine_default();
    if (!e) return COFFE_ENODEVICE;
    hipSetDevice(e->device);
    ce_spec_dev hs;
    memset(&hs, 0, sizeof(hs));
    hs.n_in = (int32_t)nin;
    ce_grid_dev hg;
    memset(&hg, 0, sizeof(hg));
    hg.kmin = kmin;
    hg.kmax = kmax;
    ce_table_desc td;
    memset(&td, 0, sizeof(td));
    td.n = n;
    td.l = l;
    const size_t off_out = 2 * nin, off_meta = off_out + nodes;
    const size_t meta = sizeof(hs) + sizeof(hg) + sizeof(td) + sizeof(int) * nodes + 64;
    pthread_mutex_lock(&e->lock);
    st = scratch_reserve(e, sizeof(double) * off_meta + meta, e->stream);
    if (st != COFFE_OK) { pthread_mutex_unlock(&e->lock); return st; }
    double *base = (double *)e->d_scratch;
    char *m = (char *)(base + off_meta);
    ce_spec_dev *d_spec = (ce_spec_dev *)m;
    ce_grid_dev *d_grid = (ce_grid_dev *)(m + sizeof(hs));
    ce_table_desc *d_tab = (ce_table_desc *)((char *)d_grid + sizeof(hg));
    int *d_stat = (int *)((char *)d_tab + sizeof(td));
    int h_stat[1 + CE_MIN_SEP_LEN];
    hipStream_t s = e->stream;
    st = hip_ok(hipMemcpyAsync(base, k, sizeof(double) * nin, hipMemcpyHostToDevice, s));
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(base + nin, pk, sizeof(double) * nin, hipMemcpyHostToDevice, s));
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(d_spec, &hs, sizeof(hs), hipMemcpyHostToDevice, s));
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(d_grid, &hg, sizeof(hg), hipMemcpyHostToDevice, s));
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(d_tab, &td, sizeof(td), hipMemcpyHostToDevice, s));
    if (st == COFFE_OK &&
        ce_launch_small_r(s, 1, 1, d_spec, d_grid, d_tab, base, base + nin, 1e-5, base + off_out, d_stat))
        st = COFFE_EDEVICE;
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(out_y, base + off_out, sizeof(double) * nodes, hipMemcpyDeviceToHost, s));
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(h_stat, d_stat, sizeof(int) * nodes, hipMemcpyDeviceToHost, s));
    if (st == COFFE_OK) st = hip_ok(hipStreamSynchronize(s));
    pthread_mutex_unlock(&e->lock);
    if (st != COFFE_OK) return st;
    st = check_small_r_status(h_stat, nodes);
    if (st != COFFE_OK) return st;
    static const double min_sep[CE_MIN_SEP_LEN] = {
        1E-6, 1E-5, 1E-4, 1E-3, 2E-3, 5E-3, 7E-3, 8E-3, 9E-3, 1E-2, 1.1E-2, 1.2E-2, 1.25E-2, 1.3E-2,
        1.35E-2, 1.5E-2, 2E-2, 2.5E-2, 3E-2, 3.5E-2, 5E-2, 7E-2, 8E-2, 9E-2, 1E-1, 1.1E-1, 1.2E-1,
        1.25E-1, 1.3E-1, 1.35E-1, 1.5E-1, 1.75E-1, 2E-1, 2.5E-1, 3E-1, 3.5E-1, 4E-1, 5E-1, 6E-1,
        7E-1, 8E-1, 9E-1};
    out_x[0] = 0.0;
    for (size_t i = 0; i < CE_MIN_SEP_LEN; ++i) out_x[1 + i] = min_sep[i];
    return COFFE_OK;
}

int coffe_integral_table(int n, int l, size_t N, const double *k, const double *pk, size_t nin, double kmin,
                         double kmax, double *out_x, double *out_y)
{
    if (!k || !pk || !out_x || !out_y || nin < 5 || n < -2 || n > 4 || l < 0 || l > 4) return COFFE_EINVAL;
    return integral_table_any((double)n, (double)l, N, k, pk, nin, kmin, kmax, out_x, out_y);
}

int coffe_integral_table_state(int n, int l, int state_n, int state_l, size_t N, const double *k, const double *pk,
                               size_t nin, double kmin, double kmax, double *out_x, double *out_y)
{
    if (state_n < 0 || state_n > 1 || state_l < 0 || state_l > 1) return COFFE_EINVAL;
    if (state_n != state_l) return COFFE_EUNSUPPORTED;
    if (state_n == 0) return coffe_integral_table(n, l, N, k, pk, nin, kmin, kmax, out_x, out_y);
    /* half-integers n/2, l/2: odd raw values, l/2 >= -1/2 (J_{l/2 + 1/2} of the small-r
       nodes has a non-negative integer order) */
    if (!k || !pk || !out_x || !out_y || nin < 5 || !(n & 1) || !(l & 1) || l < -1 || l > 9 || n < -3 || n > 9)
        return COFFE_EINVAL;
    return integral_table_any(n / 2.0, l / 2.0, N, k, pk, nin, kmin, kmax, out_x, out_y);
}

/* ---- flat-sky lensing multipoles ------------------------------------------------ */

static int flat_table_ok(int64_t off, int len, int64_t ntab_values)
{
    return len >= 5 && off >= 0 && off + len <= ntab_values;
}

/* mode 0: multipoles over ls (tables per multipole); mode 1: correlation function over
   mus (the (1/2, -1/2) table in slot 0 of ll_off / ll_len) */
static int flat_lensing_run(int mode, int nz, const double *zpar, int nl, const int *ls, const double *mus, int nsep,
                            const double *sep, int nb, const double *bg_z, const double *bg_chi, const double *bg_D1,
                            int64_t ntab_values, const double *tab_x, const double *tab_y, const int64_t *ll_off,
                            const int *ll_len, const int64_t *dl_off, const int *dl_len, double Omega_m, double H0h,
                            int want_ll, int want_dl, int panels, double *out)
{
    coffe_engine *e = coffe_engine_default();
    if (!e) return COFFE_ENODEVICE;
    hipSetDevice(e->device);
    /* the x rule: panels of 32 Gauss-Legendre nodes on [0, 1] */
    const int ngl = 32 * panels;
    double t32[64];
    gauss_legendre(32, t32, t32 + 32);
    double *gl = (double *)malloc(sizeof(double) * 2 * (size_t)ngl);
    if (!gl) return COFFE_ENOMEM;
    for (int p = 0; p < panels; ++p) {
        const double a = (double)p / panels, b = (double)(p + 1) / panels;
        for (int i = 0; i < 32; ++i) {
            gl[32 * p + i] = a + 0.5 * (b - a) * (t32[i] + 1.0);
            gl[ngl + 32 * p + i] = 0.5 * (b - a) * t32[32 + i];
        }
    }
    /* one device buffer, 8-byte slots; int arrays as int32 at the start of their slots */
    const size_t nzp = 7 * (size_t)nz, nd = (size_t)nl * CE_FL_MAXK, nout = (size_t)nz * nl * nsep;
    const size_t o_zpar = 0, o_ls = o_zpar + nzp, o_sep = o_ls + (size_t)nl, o_bz = o_sep + (size_t)nsep,
                 o_bc = o_bz + (size_t)nb, o_bd = o_bc + (size_t)nb, o_tx = o_bd + (size_t)nb,
                 o_ty = o_tx + (size_t)ntab_values, o_llo = o_ty + (size_t)ntab_values, o_lll = o_llo + (size_t)nl,
                 o_dlo = o_lll + (size_t)nl, o_dll = o_dlo + nd, o_gl = o_dll + nd, o_out = o_gl + 2 * (size_t)ngl,
                 total = o_out + nout;
    int32_t *i32 = (int32_t *)calloc(2 * (size_t)nl + nd, sizeof(int32_t));
    int64_t *i64 = (int64_t *)calloc((size_t)nl + nd, sizeof(int64_t));
    if (!i32 || !i64) {
        free(gl);
        free(i32);
        free(i64);
        return COFFE_ENOMEM;
    }
    int32_t *ls32 = i32, *lll32 = i32 + nl, *dll32 = i32 + 2 * (size_t)nl;
    int64_t *llo64 = i64, *dlo64 = i64 + nl;
    for (int il = 0; il < nl; ++il) {
        ls32[il] = mode == 0 ? ls[il] : 0;
        lll32[il] = ll_len[mode == 0 ? il : 0];
        llo64[il] = ll_off[mode == 0 ? il : 0];
    }
    if (mode == 0)
        for (size_t t = 0; t < nd; ++t) {
            dll32[t] = dl_len[t];
            dlo64[t] = dl_off[t];
        }
    double *d = NULL;
    int st = dev_alloc((void **)&d, sizeof(double) * total);
    hipStream_t s = e->stream;
#define FL_PUT(off, src, bytes) \
    if (st == COFFE_OK && (bytes) > 0) st = hip_ok(hipMemcpyAsync(d + (off), (src), (bytes), hipMemcpyHostToDevice, s))
    FL_PUT(o_zpar, zpar, sizeof(double) * nzp);
    if (mode == 0) {
        FL_PUT(o_ls, ls32, sizeof(int32_t) * (size_t)nl);
    } else {
        FL_PUT(o_ls, mus, sizeof(double) * (size_t)nl);
    }
    FL_PUT(o_sep, sep, sizeof(double) * (size_t)nsep);
    FL_PUT(o_bz, bg_z, sizeof(double) * (size_t)nb);
    FL_PUT(o_bc, bg_chi, sizeof(double) * (size_t)nb);
    FL_PUT(o_bd, bg_D1, sizeof(double) * (size_t)nb);
    FL_PUT(o_tx, tab_x, sizeof(double) * (size_t)ntab_values);
    FL_PUT(o_ty, tab_y, sizeof(double) * (size_t)ntab_values);
    FL_PUT(o_llo, llo64, sizeof(int64_t) * (size_t)nl);
    FL_PUT(o_lll, lll32, sizeof(int32_t) * (size_t)nl);
    FL_PUT(o_dlo, dlo64, sizeof(int64_t) * nd);
    FL_PUT(o_dll, dll32, sizeof(int32_t) * nd);
    FL_PUT(o_gl, gl, sizeof(double) * 2 * (size_t)ngl);
#undef FL_PUT
    if (st == COFFE_OK &&
        ce_launch_flat_lensing(s, mode, d + o_ls, nz, nl, nsep, nb, ngl, want_ll != 0, want_dl != 0, d + o_zpar,
                               (const int *)(d + o_ls), d + o_sep, d + o_bz, d + o_bc, d + o_bd, d + o_tx, d + o_ty,
                               (const int64_t *)(d + o_llo), (const int *)(d + o_lll), (const int64_t *)(d + o_dlo),
                               (const int *)(d + o_dll), d + o_gl, d + o_gl + ngl, Omega_m, H0h, d + o_out))
        st = COFFE_EDEVICE;
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(out, d + o_out, sizeof(double) * nout, hipMemcpyDeviceToHost, s));
    if (st == COFFE_OK) st = hip_ok(hipStreamSynchronize(s));
    hipFree(d);
    free(i32);
    free(i64);
    free(gl);
    return st;
}

static int flat_common_ok(int nz, const double *zpar, int nl, int nsep, const double *sep, int nb, const double *bg_z,
                          const double *bg_chi, const double *bg_D1, int64_t ntab_values, const double *tab_x,
                          const double *tab_y, int panels, const double *out)
{
    return !(nz <= 0 || nl <= 0 || nsep <= 0 || nb < 5 || !zpar || !sep || !bg_z || !bg_chi || !bg_D1 || !out ||
             panels < 1 || panels > 4096 || ntab_values < 0 || (ntab_values > 0 && (!tab_x || !tab_y)) ||
             (int64_t)nz * nl * nsep > 0x7fffffff);
}

int coffe_flat_lensing_multipoles(int nz, const double *zpar, int nl, const int *ls, int nsep, const double *sep,
                                  int nb, const double *bg_z, const double *bg_chi, const double *bg_D1,
                                  int64_t ntab_values, const double *tab_x, const double *tab_y, const int64_t *ll_off,
                                  const int *ll_len, const int64_t *dl_off, const int *dl_len, double Omega_m,
                                  double H0h, int want_ll, int want_dl, int panels, double *out)
{
    if (!flat_common_ok(nz, zpar, nl, nsep, sep, nb, bg_z, bg_chi, bg_D1, ntab_values, tab_x, tab_y, panels, out) ||
        !ls || !ll_off || !ll_len || !dl_off || !dl_len)
        return COFFE_EINVAL;
    /* every table a wanted term reads must be there; the others must be empty or whole */
    for (int il = 0; il < nl; ++il) {
        const int l = ls[il];
        if (l < 0) return COFFE_EINVAL;
        const int even = (l % 2) == 0;
        if (want_ll && even && !flat_table_ok(ll_off[il], ll_len[il], ntab_values)) return COFFE_EINVAL;
        if (ll_len[il] != 0 && !flat_table_ok(ll_off[il], ll_len[il], ntab_values)) return COFFE_EINVAL;
        if (want_dl && even && l > 2 * (CE_FL_MAXK - 1)) return COFFE_EINVAL;
        for (int k = 0; k < CE_FL_MAXK; ++k) {
            const int t = il * CE_FL_MAXK + k;
            if (want_dl && even && k <= l / 2 && !flat_table_ok(dl_off[t], dl_len[t], ntab_values))
                return COFFE_EINVAL;
            if (dl_len[t] != 0 && !flat_table_ok(dl_off[t], dl_len[t], ntab_values)) return COFFE_EINVAL;
        }
    }
    return flat_lensing_run(0, nz, zpar, nl, ls, NULL, nsep, sep, nb, bg_z, bg_chi, bg_D1, ntab_values, tab_x, tab_y,
                            ll_off, ll_len, dl_off, dl_len, Omega_m, H0h, want_ll, want_dl, panels, out);
}

int coffe_flat_lensing_corrfunc(int nz, const double *zpar, int nmu, const double *mu, int nsep, const double *sep,
                                int nb, const double *bg_z, const double *bg_chi, const double *bg_D1, int ntab,
                                const double *tab_x, const double *tab_y, double Omega_m, double H0h, int want_ll,
                                int want_dl, int panels, double *out)
{
    if (!flat_common_ok(nz, zpar, nmu, nsep, sep, nb, bg_z, bg_chi, bg_D1, ntab, tab_x, tab_y, panels, out) || !mu ||
        ((want_ll || want_dl) && ntab < 5))
        return COFFE_EINVAL;
    for (int i = 0; i < nmu; ++i)
        if (!(mu[i] >= -1.0 && mu[i] <= 1.0)) return COFFE_EINVAL;
    const int64_t off0 = 0;
    const int len0 = ntab;
    return flat_lensing_run(1, nz, zpar, nmu, NULL, mu, nsep, sep, nb, bg_z, bg_chi, bg_D1, ntab, tab_x, tab_y, &off0,
                            &len0, NULL, NULL, Omega_m, H0h, want_ll, want_dl, panels, out);
}

/* ---- integrated terms, full sky ---------------------------------------------- */

/* the table slots of the integrated kernel, in the order of integrated_kernels.hip */
static const int intg_slot_n[CE_INTG_NSLOT] = {0, 0, 0, 1, 1, 2, 2, 3, 4};
static const int intg_slot_l[CE_INTG_NSLOT] = {0, 2, 4, 1, 3, 0, 2, 1, 0};

/* the slots the terms of `flags` read (bit t: slot t) */
static unsigned intg_needed_slots(unsigned f)
{
    const unsigned loc = COFFE_INTEGRATED_DEN | COFFE_INTEGRATED_RSD | COFFE_INTEGRATED_D1 | COFFE_INTEGRATED_D2 |
                         COFFE_INTEGRATED_G1 | COFFE_INTEGRATED_G2 | COFFE_INTEGRATED_G3;
    const int len = (f & COFFE_INTEGRATED_LEN) != 0, g45 = (f & (COFFE_INTEGRATED_G4 | COFFE_INTEGRATED_G5)) != 0;
    const int flat_dl = (f & COFFE_INTEGRATED_FLAT_DL) != 0;
    unsigned m = 0;
    enum { S00, S02, S04, S11, S13, S20, S22, S31, S40 };
#define B(s) (1u << (s))
    if (len && !(f & COFFE_INTEGRATED_ONLY_CROSS) && !(f & COFFE_INTEGRATED_FLAT_LL))
        m |= B(S00) | B(S02) | B(S04) | B(S11) | B(S13) | B(S20);
    if (len && g45) m |= B(S31) | B(S22);
    if (g45) m |= B(S40);
    if (len && !flat_dl && (f & COFFE_INTEGRATED_DEN)) m |= B(S11) | B(S02);
    if (len && !flat_dl && (f & COFFE_INTEGRATED_RSD)) m |= B(S00) | B(S02) | B(S04);
    if (len && !flat_dl && (f & COFFE_INTEGRATED_D1)) m |= B(S11) | B(S20) | B(S13);
    if (len && (f & (COFFE_INTEGRATED_D2 | COFFE_INTEGRATED_G1 | COFFE_INTEGRATED_G2 | COFFE_INTEGRATED_G3)))
        m |= B(S31) | B(S22);
    if (g45 && (f & COFFE_INTEGRATED_DEN)) m |= B(S20);
    if (g45 && (f & COFFE_INTEGRATED_RSD)) m |= B(S22) | B(S20);
    if (g45 && (f & COFFE_INTEGRATED_D1)) m |= B(S31);
    if (g45 && (f & loc)) m |= B(S40);
#undef B
    return m;
}

int coffe_integrated_corrfunc(const coffe_mu_inputs *in, unsigned flags, int ntab, const int *tab_n, const int *tab_l,
                              const int *tab_len, const double *tab_x, const double *tab_y, int nren0,
                              const double *ren0_x, const double *ren0_y, int nz, const double *zmean, int nmu,
                              const double *mu, int nsep, const double *sep, const int *rule, double *out)
{
    const unsigned all = (COFFE_INTEGRATED_FLAT_DL << 1) - 1;
    if (!in || (flags & ~all) || nz <= 0 || nmu <= 0 || nsep <= 0 || !zmean || !mu || !sep || !out || ntab < 0 ||
        ntab > 64 || (ntab > 0 && (!tab_n || !tab_l || !tab_len || !tab_x || !tab_y)) ||
        (int64_t)nz * nmu * nsep > 0x7fffffff)
        return COFFE_EINVAL;
    if (in->nbg < 5 || !in->bg_z || !in->bg_chi || !in->bg_D1 || !in->bg_a || !in->bg_f || !in->bg_cH ||
        !in->bg_G1 || !in->bg_G2 || in->nb < 5 || !in->bias_z || !in->bias1 || !in->bias2 || in->nmb < 5 ||
        !in->mb_z || !in->s1 || !in->s2 || !in->fevo1 || !in->fevo2)
        return COFFE_EINVAL;
    int r[6] = {8, 16, 12, 16, 32, 16};
    if (rule)
        for (int i = 0; i < 6; ++i) r[i] = rule[i];
    for (int i = 0; i < 6; i += 2)
        if (r[i] < 1 || r[i] > 256 || r[i + 1] < 2 || r[i + 1] > 64) return COFFE_EINVAL;
    if ((r[0] + (r[0] / 2 > 1 ? r[0] / 2 : 1)) * r[1] > CE_INTG_MAX_X1) return COFFE_EINVAL;
    for (int i = 0; i < nmu; ++i)
        if (!(mu[i] >= -1.0 && mu[i] <= 1.0)) return COFFE_EINVAL;
    for (int i = 0; i < nsep; ++i)
        if (!(sep[i] > 0.0) || !isfinite(sep[i])) return COFFE_EINVAL;
    /* along the line of sight (|mu| -> 1) the separations between the integration points
       reach FP64 rounding noise and the terms' r = 0 branches; 1 - 1e-10 keeps the
       closest approach above it */
    for (int i = 0; i < nmu; ++i)
        if (fabs(mu[i]) > 1.0 - 1e-10) return COFFE_EDOMAIN;
    /* the tables by slot; a needed one must be there with 5 nodes or more */
    int off[CE_INTG_NSLOT], len[CE_INTG_NSLOT];
    int64_t total_tab = 0;
    for (int t = 0; t < CE_INTG_NSLOT; ++t) off[t] = len[t] = 0;
    for (int i = 0; i < ntab; ++i) {
        if (tab_len[i] < 0 || total_tab + tab_len[i] > 0x7fffffff) return COFFE_EINVAL;
        for (int t = 0; t < CE_INTG_NSLOT; ++t)
            if (tab_n[i] == intg_slot_n[t] && tab_l[i] == intg_slot_l[t]) {
                off[t] = (int)total_tab;
                len[t] = tab_len[i];
            }
        total_tab += tab_len[i];
    }
    const unsigned need = intg_needed_slots(flags);
    for (int t = 0; t < CE_INTG_NSLOT; ++t)
        if ((need >> t & 1u) && len[t] < 5) return COFFE_EINVAL;
    const int want_ren = (need >> 8) & 1u;
    if (want_ren && (in->nren < 5 || !in->ren_chi || !in->ren || nren0 < 5 || !ren0_x || !ren0_y))
        return COFFE_EINVAL;
    coffe_engine *e = coffe_engine_default();
    if (!e) return COFFE_ENODEVICE;
    hipSetDevice(e->device);
    const size_t nbg = (size_t)in->nbg, nb = (size_t)in->nb, nmb = (size_t)in->nmb;
    const size_t nren = want_ren ? (size_t)in->nren : 0, nnode = nren * nren, n0 = want_ren ? (size_t)nren0 : 0;
    const size_t ngl = 2 * (size_t)(r[1] + r[3] + r[5]), nout = 2 * (size_t)nz * nmu * nsep;
    const int nguess = 4096;
    /* the Akima value records and node pairs, and 7 doubles of slack to align them to
       64 bytes */
    const size_t nrec = (CE_INTG_VAL + CE_INTG_PAIR) * ((size_t)total_tab + CE_INTG_NBGC * nbg) + 7;
    const size_t total = 8 * nbg + 3 * nb + 5 * nmb + 2 * (size_t)total_tab + 2 * n0 + nren + 4 * nnode +
                         (size_t)nz + (size_t)nmu + (size_t)nsep + ngl + (size_t)nguess + 2 + nout + nrec +
                         CE_INTG_ARGS_BYTES / sizeof(double);
    double *h = (double *)malloc(sizeof(double) * (total - nout - nrec - CE_INTG_ARGS_BYTES / sizeof(double)));
    double *work = (double *)malloc(sizeof(double) * (2 * nren + 2));
    double *planes = (double *)malloc(sizeof(double) * (4 * nnode + 1));
    if (!h || !work || !planes) {
        free(h);
        free(work);
        free(planes);
        return COFFE_ENOMEM;
    }
    double *d = NULL;
    int st = dev_alloc((void **)&d, sizeof(double) * total);
    ce_intg_dev dv;
    memset(&dv, 0, sizeof(dv));
    size_t o = 0;
    /* place n values of src in the staging buffer; the device pointer at the same offset */
#define IG_PUT(src, n)                                                       \
    ({                                                                       \
        const size_t n_ = (size_t)(n);                                       \
        if (n_) memcpy(h + o, (src), sizeof(double) * n_);                   \
        const double *r_ = d + o;                                            \
        o += n_;                                                             \
        r_;                                                                  \
    })
    dv.nbg = in->nbg;
    dv.bg_z = IG_PUT(in->bg_z, nbg);
    dv.bg_chi = IG_PUT(in->bg_chi, nbg);
    dv.bg_D1 = IG_PUT(in->bg_D1, nbg);
    dv.bg_a = IG_PUT(in->bg_a, nbg);
    dv.bg_f = IG_PUT(in->bg_f, nbg);
    dv.bg_cH = IG_PUT(in->bg_cH, nbg);
    dv.bg_G1 = IG_PUT(in->bg_G1, nbg);
    dv.bg_G2 = IG_PUT(in->bg_G2, nbg);
    dv.nb = in->nb;
    dv.bias_z = IG_PUT(in->bias_z, nb);
    dv.b1 = IG_PUT(in->bias1, nb);
    dv.b2 = IG_PUT(in->bias2, nb);
    dv.nmb = in->nmb;
    dv.mb_z = IG_PUT(in->mb_z, nmb);
    dv.s1 = IG_PUT(in->s1, nmb);
    dv.s2 = IG_PUT(in->s2, nmb);
    dv.fe1 = IG_PUT(in->fevo1, nmb);
    dv.fe2 = IG_PUT(in->fevo2, nmb);
    dv.tab_x = IG_PUT(tab_x, total_tab);
    dv.tab_y = IG_PUT(tab_y, total_tab);
    for (int t = 0; t < CE_INTG_NSLOT; ++t) {
        dv.tab_off[t] = off[t];
        dv.tab_len[t] = len[t];
        /* the tables of coffe_integral_table: r = 0, the 42 small separations, then N
           log-uniform FFT nodes from index 43 */
        const double *x = tab_x + off[t];
        const int n = len[t], f0 = 1 + CE_MIN_SEP_LEN;
        dv.tab_f0[t] = -1;
        if (n > f0 + 4 && x[0] == 0.0 && x[f0] > 0.0 && x[n - 1] > x[f0]) {
            dv.tab_f0[t] = f0;
            dv.tab_lx0[t] = log(x[f0]);
            dv.tab_lsc[t] = (double)(n - 1 - f0) / (log(x[n - 1]) - log(x[f0]));
        }
    }
    /* one interval serves every standard table when they share their nodes (the tables of
       one spectrum and one N do) */
    dv.tab_ref = -1;
    for (int t = 0; t < CE_INTG_NSLOT - 1; ++t) {
        if (len[t] < 5) continue;
        if (dv.tab_ref < 0) {
            dv.tab_ref = t;
        } else if (len[t] != len[dv.tab_ref] ||
                   memcmp(tab_x + off[t], tab_x + off[dv.tab_ref], sizeof(double) * (size_t)len[t]) != 0) {
            dv.tab_ref = -1;
            break;
        }
    }
    /* guesses of bg_chi's interval on a uniform chi grid */
    {
        double *g = h + o;
        const double cmax = in->bg_chi[nbg - 1];
        size_t j = 0;
        for (int k = 0; k <= nguess; ++k) {
            const double c = cmax * k / nguess;
            while (j + 2 < nbg && in->bg_chi[j + 1] <= c) ++j;
            g[k] = (double)j;
        }
        dv.nguess = nguess;
        dv.chi_guess = d + o;
        o += (size_t)nguess + 1;
    }
    if (want_ren) {
        dv.nren0 = nren0;
        dv.ren0_x = IG_PUT(ren0_x, n0);
        dv.ren0_y = IG_PUT(ren0_y, n0);
        dv.nren = in->nren;
        dv.ren_chi = IG_PUT(in->ren_chi, nren);
        /* values and the bicubic's node-derivative tables (as coffe_mu_weights_ex),
           interleaved per node: one node's four numbers are two 16-byte loads */
        double *hv = planes, *hx = hv + nnode, *hy = hx + nnode, *hxy = hy + nnode;
        memcpy(hv, in->ren, sizeof(double) * nnode);
        for (size_t j = 0; j < nren; ++j) spline_node_slopes((int)nren, in->ren_chi, hv + j * nren, 1, hx + j * nren, 1, work);
        for (size_t i = 0; i < nren; ++i) spline_node_slopes((int)nren, in->ren_chi, hv + i, (int)nren, hy + i, (int)nren, work);
        for (size_t i = 0; i < nren; ++i) spline_node_slopes((int)nren, in->ren_chi, hx + i, (int)nren, hxy + i, (int)nren, work);
        o += o & 1;
        for (size_t k = 0; k < nnode; ++k) {
            h[o + 4 * k] = hv[k];
            h[o + 4 * k + 1] = hx[k];
            h[o + 4 * k + 2] = hy[k];
            h[o + 4 * k + 3] = hxy[k];
        }
        dv.ren_rec = d + o;
        o += 4 * nnode;
    }
    dv.Om = in->Omega_m;
    dv.H0h = in->H0h;
    const double *d_z = IG_PUT(zmean, nz);
    const double *d_mu = IG_PUT(mu, nmu);
    const double *d_sep = IG_PUT(sep, nsep);
    const double *d_gl = d + o;
    {
        const int ns[3] = {r[1], r[3], r[5]};
        for (int k = 0; k < 3; ++k) {
            gauss_legendre(ns[k], h + o, h + o + ns[k]);
            o += 2 * (size_t)ns[k];
        }
    }
#undef IG_PUT
    double *d_out = d + o;
    dv.ntab_values = (int)total_tab;
    dv.coef = (double *)(((uintptr_t)(d_out + nout) + 63) & ~(uintptr_t)63);
    dv.args_dev = d_out + nout + nrec;
    hipStream_t s = e->stream;
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(d, h, sizeof(double) * o, hipMemcpyHostToDevice, s));
    if (st == COFFE_OK && ce_launch_integrated(s, &dv, flags, nz, d_z, nmu, d_mu, nsep, d_sep, r, d_gl, d_out))
        st = COFFE_EDEVICE;
    if (st == COFFE_OK) st = hip_ok(hipMemcpyAsync(out, d_out, sizeof(double) * nout, hipMemcpyDeviceToHost, s));
    if (st == COFFE_OK) st = hip_ok(hipStreamSynchronize(s));
    else if (d) hipStreamSynchronize(s);
    hipFree(d);
    free(h);
    free(work);
    free(planes);
    return st;
}
