/*
 * Internal interface between the host C layer (engine.c, dropin.c) and the HIP
 * kernels (fftlog_kernels.hip, mu_kernels.hip, quad_kernels.hip).  Plain C types only.
 */
#ifndef COFFE_ENGINE_INTERNAL_H
#define COFFE_ENGINE_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CE_MAX_PASSES 16
#define CE_MAX_TABLES 8
#define CE_MIN_SEP_LEN 42
/* half-integer tables per multipole of the flat-sky density-lensing term: k <= l/2,
   so multipoles up to l = 2 (CE_FL_MAXK - 1) */
#define CE_FL_MAXK 5
/* intervals an adaptive QAG keeps in LDS (the reference allows 50000; the small-r
   integrands converge within a few hundred) */
#define CE_QAG_CAP 1024
/* input intervals per spline chunk and LDS nodes staged per chunk (chunk + 2 halos
   of 64 + end nodes) */
#define CE_CHUNK 512
#define CE_CHUNK_NODES 648

/* FFT radix schedule for a complex transform of length M = N/2 */
/* complex LDS slots of one transform of length M (the swizzled layout of the FFT
   kernels permutes slots inside aligned blocks of eight) */
#define CE_LDS_DATA_LEN(M) ((((M) + 7) / 8) * 8)

typedef struct ce_radix_plan {
    int npass;
    int radix[CE_MAX_PASSES];
    /* the same schedule packed 3 bits a pass, first pass lowest, 0 ends it: codes
       1..5 = radix 2, 3, 4, 5, 8 and 7 = radix 10 (at most 10 passes) */
    uint32_t code;
} ce_radix_plan;

/* spectrum-independent tables for one (N, kmin, kmax) grid, device pointers */
typedef struct ce_grid_dev {
    ce_radix_plan rp;  /* radix schedule of the length-M complex FFT */
    int N;             /* FFT length (even) */
    int M;             /* N / 2 */
    double kmin, kmax, G;
    double k0, r0;     /* x_i = k0 p_i, r_i = r0 p_i */
    const double *p;   /* (kmax/kmin)^(i/N), i < N  (x_i = kmin p_i, r_i = p_i) */
    const double *tw;  /* interleaved e^{-2 pi i k/N}, k < N */
    const double *bw[2];  /* bias (kmax/kmin)^((3-q')i/N) times the k-window, per distinct q' */
    const double *U[CE_MAX_TABLES];   /* interleaved: kernel u_l(t_k) * W_F(k) / L, k <= M */
    const double *os[CE_MAX_TABLES];  /* output scale: prefactor_i * r_i^(n-l), i < N */
} ce_grid_dev;

/* one spectrum of a batch (host side view, copied to device as an array) */
typedef struct ce_spec_dev {
    int64_t in_off;    /* offset of k / P in the concatenated input */
    int32_t n_in;      /* input length */
    int32_t grid;      /* index into the grid array */
    int64_t xf_off;    /* offset (complex units) of the r2c workspace: nq * (M+1) */
    int64_t xb_off;    /* offset (complex units) of the adjoint workspace: ntab * (M+1) */
    int64_t ws_off;    /* offset (doubles) of the real workspace: 2 N (S(x_i), then its cotangent) */
} ce_spec_dev;

/* one chunk of input intervals [a, b) of a spectrum (b = n_in - 1 on the last) */
typedef struct ce_chunk_dev {
    int32_t spec;
    int32_t a, b;
    int32_t grid;      /* copies of the spectrum's fields: one load gives everything */
    int64_t in_off, ws_off;
    int32_t n_in;
    int32_t id;        /* the chunk's index in the batch: its slot of the resample ranges */
} ce_chunk_dev;

/* work unit of the table kernels */
typedef struct ce_unit_dev {
    int32_t spec;
    int32_t tab;       /* table slot */
} ce_unit_dev;

typedef struct ce_table_desc {
    double n, l;       /* power nu and Bessel order of the transform */
    int qidx;          /* which distinct q' = q + nu (0/1) */
    int rpow;          /* multiply by r^(n-l) when n >= l (the I_l^n tables) */
} ce_table_desc;

/* ---- launchers (fftlog_kernels.hip); all return 0 or a HIP error code ---- */
int ce_launch_grid_tables(void *stream, int N, double kmin, double kmax, double k0, double r0, const double *d_p,
                          int ntab, const ce_table_desc *tabs, const double *tab_qnu, int nq, const double *qprime,
                          double *const *d_U, double *const *d_os, double *const *d_bw);

/* d_packs (4 ints per workgroup: first unit, unit count, groups, 0; NULL: one unit
   per workgroup): units of one transform length share a workgroup, one group of
   threads / groups threads and lds_data_len(M) complex of LDS each; lt_off = LDS
   offset of the twiddle tables (complex), at least groups * lds_data_len(M) of every
   pack (0: lds_data_len(max_M)) */
/* d_ranges (4 ints per chunk id, may be NULL): the resample index ranges of the
   chunk's owned and staged intervals, written for ce_launch_spline_adj */
/* flat-sky lensing multipoles (lensing_kernels.hip), device pointers */
int ce_launch_flat_lensing(void *stream, int mode, const double *d_mus, int nz, int nl, int nsep, int nb, int ngl,
                           int want_ll, int want_dl, const double *d_zpar, const int *d_ls, const double *d_sep,
                           const double *d_bg_z,
                           const double *d_bg_chi, const double *d_bg_D1, const double *d_tab_x, const double *d_tab_y,
                           const int64_t *d_ll_off, const int *d_ll_len, const int64_t *d_dl_off, const int *d_dl_len,
                           const double *d_glx, const double *d_glw, double Omega_m, double H0h, double *d_out);
/* complex FFTs of any length on device memory (anyfft.hip): forward, unnormalised,
   in place on d_x (L complex values); direct LDS lines up to 8192 points, four-step for
   longer 2^a 3^b 5^c lengths, Bluestein otherwise */
#define CE_ANY_MAX_N (1 << 21)
typedef struct ce_anyfft ce_anyfft;
int ce_anyfft_create(int64_t L, ce_anyfft **out);
void ce_anyfft_destroy(ce_anyfft *p);
int ce_anyfft_forward(ce_anyfft *p, void *d_x, void *stream);
/* nb transforms of L points one after another (nb <= 65535) */
int ce_anyfft_forward_batch(ce_anyfft *p, void *d_x, int64_t nb, void *stream);
int ce_anyfft_kind(const ce_anyfft *p);
int64_t ce_anyfft_padded(const ce_anyfft *p);
/* the FFTLog stages around the two length-N transforms of the general path */
int ce_launch_tf_in(void *stream, const double *d_S, const double *d_bw, void *d_c, int N);
int ce_launch_tf_mid(void *stream, void *d_c, const double *d_U, int N);
int ce_launch_tf_out(void *stream, const void *d_c, const double *d_os, double *d_out, int N);
/* the general-length counterparts of ce_launch_cov_ahat and ce_launch_cov_rows (same
   arrays; the row transform synchronises the stream) */
int ce_general_cov_ahat(ce_anyfft *fft, void *stream, int nl, int N, const double *d_A, double *d_ahat);
int ce_general_cov_rows(ce_anyfft *fft, void *stream, int nblk, int N, const double *d_ahat, const double *d_B,
                        const double *d_d, int npair, const int *d_pl1, const int *d_pl2, int nrows,
                        const int *d_rows, const double *d_prn1, int R, const int *d_flo, const int *d_fcnt,
                        const int *d_foff, const double *d_fw, double *d_W);

int ce_launch_resample(void *stream, int nchunk, const ce_chunk_dev *d_chunks, const ce_spec_dev *d_spec,
                       const ce_grid_dev *d_grids, const double *d_k, const double *d_pk, double *d_ws,
                       int32_t *d_ranges);

int ce_launch_prepare(void *stream, int nspec, const ce_spec_dev *d_spec, const ce_grid_dev *d_grids, int nq,
                      const double *d_ws, double *d_xf, int threads, int max_M);

int ce_launch_tables(void *stream, int nunit, const ce_unit_dev *d_units, const ce_spec_dev *d_spec,
                     const ce_grid_dev *d_grids, const ce_table_desc *d_tabs, int ntab, const double *d_xf,
                     int nsep, const int32_t *d_sidx, const double *d_sep, double *d_itab, double *d_full,
                     int64_t full_stride, const double *d_A, const double *d_gout, int nz, int nl, double *d_xbar,
                     double *d_wbuf, int threads, int max_M, const int32_t *d_packs, int npack, int lt_off);

int ce_launch_project(void *stream, int nspec, int ntab, int nz, int nl, int nsep, const double *d_A,
                      const double *d_itab, double *d_out);

/* with d_out (and d_A, d_itab) each spectrum's workgroup also writes its multipoles
   out[s][z][l][r] = sum_t A itab (the projection of ce_launch_project) */
int ce_launch_adj_c2r(void *stream, int nspec, const ce_spec_dev *d_spec, const ce_grid_dev *d_grids,
                      const ce_table_desc *d_tabs, int ntab, int nq, double *d_ws, const double *d_xbar, int threads,
                      int max_M, const double *d_A, const double *d_itab, double *d_out, int nz, int nl, int nsep);

int ce_launch_spline_adj(void *stream, int nchunk, const ce_chunk_dev *d_chunks, const ce_spec_dev *d_spec,
                         const ce_grid_dev *d_grids, const double *d_k, const double *d_ws, double *d_gpk,
                         const int32_t *d_ranges);

/* ---- small-separation table nodes (quad_kernels.hip): out/status [nspec][ntab][43] ---- */
int ce_launch_small_r(void *stream, int nspec, int ntab, const ce_spec_dev *d_spec, const ce_grid_dev *d_grids,
                      const ce_table_desc *d_tabs, const double *d_k, const double *d_pk, double epsrel,
                      double *d_out, int *d_status);
/* the divergent I_0^4 table and its renormalisations (quad_kernels.hip): status
   and interval counts (may be null) [2 nsep + nchi^2] in the order i40, r0, ren */
int ce_launch_divergent(void *stream, const double *d_k, const double *d_pk, int n_in, double kmin, double kmax,
                        const double *d_sep, int nsep, const double *d_chi, int nchi, double epsrel, double *d_i40,
                        double *d_r0, double *d_ren, int *d_status, int *d_nint);

/* ---- 2D FFTLog covariance (cov_kernels.hip) ---- */
size_t ce_cov_lds_bytes(int N);
int ce_launch_cov_ahat(void *stream, int nl, int N, const ce_radix_plan *rp, const double *d_tw, const double *d_A,
                       double *d_ahat);
int ce_launch_cov_diag(void *stream, int nspec, int N, const int64_t *d_in_off, const int32_t *d_n_in,
                       const double *d_k, const double *d_pk, const double *d_kg, const double *d_k3,
                       const double *d_kn1, const double *d_kn2, double step, double *d_d, int *d_status);
int ce_launch_cov_rows(void *stream, int nblk, int N, const ce_radix_plan *rpN, const ce_radix_plan *rpM,
                       const double *d_tw, const double *d_ahat, const double *d_B, const double *d_d, int npair,
                       const int *d_pl1, const int *d_pl2, int nrows, const int *d_rows, const double *d_prn1,
                       const double *d_prn2, int R, const int *d_flo, const int *d_fcnt, const int *d_foff,
                       const double *d_fw, double *d_W);
/* the adjoint of coffe_cov_integrals: d_gint -> d_gpk (zeroed here), workspaces
   d_Wbar [2 nspec][npair][nrows][R] and d_dbar [nspec][2][N] */
int ce_launch_cov_vjp(void *stream, int nspec, int N, const ce_radix_plan *rpN, const ce_radix_plan *rpM,
                      const double *d_tw, const double *d_ahat, const double *d_B, int npair, const int *d_pl1,
                      const int *d_pl2, int nrows, const int *d_rows, const double *d_prn1, int R, const int *d_flo,
                      const int *d_fcnt, const int *d_foff, const double *d_fws, const int *d_upos,
                      const int *d_uelo, const int *d_uehi,
                      const double *d_fw, const double *d_fac, const double *d_gint, double *d_Wbar, double *d_dbar,
                      const int64_t *d_in_off, const int32_t *d_n_in, const double *d_k, const double *d_pk,
                      const double *d_kg, const double *d_k3, const double *d_kn1, const double *d_kn2, double step,
                      double *d_gpk, int64_t n_total);
int ce_launch_cov_cols(void *stream, int nspec, int npair, int nrows, int R, const int *d_upos, const int *d_ucnt,
                       const int *d_uoff, const double *d_uw, const double *d_fac, const double *d_W, double *d_out);
int ce_launch_cov_assemble(void *stream, int nspec, int nz, int nl, int R, int npair, const int *d_ls,
                           const int *d_pair_of, const double *d_sep, const double *d_cc, const double *d_cm,
                           const double *d_D1, const double *d_pscale, const double *d_vol, const double *d_pix,
                           int pd_same, int pd_swap, int window, int mixed, int cosmic, const double *d_integ,
                           double *d_out);
int ce_launch_cov_assemble_vjp(void *stream, int nspec, int nz, int nl, int R, int npair, const int *d_ls,
                               const int *d_pair_of, const double *d_cc, const double *d_cm, const double *d_D1,
                               const double *d_vol, int mixed, int cosmic, const double *d_integ,
                               const double *d_gout, double *d_ginteg, double *d_gcoef);

/* ---- mu-quadrature weights (mu_kernels.hip) ---- */
/* device inputs: background tables on bg_z (a: scale factor, d2 / g terms only), galaxy
   biases on bias_z, magnification and evolution biases on mb_z (d2 / g terms), and the
   renormalisation grid of I_0^4 with the bicubic's node derivatives ([b nren + a] at
   (ren_chi[a], ren_chi[b]); d2 / g auto terms) */
typedef struct ce_mu_dev {
    int nbg;
    const double *bg_z, *bg_chi, *bg_D1, *bg_f, *bg_cH, *bg_G1, *bg_G2, *bg_a;
    int nb;
    const double *bias_z, *bias1, *bias2;
    int nmb;
    const double *mb_z, *s1, *s2, *fevo1, *fevo2;
    double H0h, Om;
    int nren;
    const double *ren_chi, *ren_v, *ren_x, *ren_y, *ren_xy;
} ce_mu_dev;
/* coef: workspace of 3 (7 nbg + 2 nb + 4 nmb) doubles for the Akima coefficients; dg: the
   d2 / g sources (bit 0 d2, 1 g1, 2 g2, 3 g3); paired: point i = (zmean[i], sep[i]) with
   nz == nsep, A [nz][nl][ntab] and ren_out [nz][nl]; else A [nz][nl][ntab][nsep] and
   ren_out [nz][nl][nsep] (ren_out may be NULL) */
int ce_launch_mu_weights(void *stream, const ce_mu_dev *in, unsigned terms, unsigned dg, int flatsky, int paired,
                         int nz, const double *zmean, int nl, const int *ls, int ntab, const int *slot, int nsep,
                         const double *sep, int nmu, const double *mu, const double *wmu, double *A, double *ren_out,
                         double *coef);
/* the tangents of the background tables along one direction (device, on the z grid) */
typedef struct ce_mu_tan {
    const double *chi, *f, *cH, *G1, *G2, *D1;
} ce_mu_tan;
/* the weights and their tangents At (layout of A) along tan, density / RSD / Doppler-1
   terms only (no d2 / g); coef: 3 (7 nbg + 2 nb) + 18 nbg doubles */
int ce_launch_mu_weights_jvp(void *stream, const ce_mu_dev *in, const ce_mu_tan *tan, unsigned terms, int flatsky,
                             int paired, int nz, const double *zmean, int nl, const int *ls, int ntab, const int *slot,
                             int nsep, const double *sep, int nmu, const double *mu, const double *wmu, double *A,
                             double *At, double *coef);

/* ---- integrated terms (integrated_kernels.hip) ---- */
/* device inputs: the background on bg_z, galaxy biases on bias_z, magnification and
   evolution biases on mb_z, the tables concatenated (slots (0,0), (0,2), (0,4), (1,1),
   (1,3), (2,0), (2,2), (3,1), (4,0): offset and length, 0 when absent), the r = 0
   renormalisation of I_0^4 (ren0) and the renormalisation grid with its bicubic's node
   derivatives ([b nren + a] at (ren_chi[a], ren_chi[b])) */
#define CE_INTG_NSLOT 9
#define CE_INTG_ARGS_BYTES 2048
#define CE_INTG_MAX_X1 512    /* x1 nodes of the double rule: (p1 + p1 / 2) n1 at most */
#define CE_INTG_VAL 4         /* doubles per Akima value record {y, b, c, d} */
#define CE_INTG_PAIR 2        /* doubles per node pair {x_j, x_(j+1)} */
#define CE_INTG_NBGC 5        /* background curves with value records */
typedef struct ce_intg_dev {
    int nbg;
    const double *bg_z, *bg_chi, *bg_D1, *bg_a, *bg_f, *bg_cH, *bg_G1, *bg_G2;
    int nb;
    const double *bias_z, *b1, *b2;
    int nmb;
    const double *mb_z, *s1, *s2, *fe1, *fe2;
    const double *tab_x, *tab_y;
    int tab_off[CE_INTG_NSLOT], tab_len[CE_INTG_NSLOT];
    /* lookup guesses: the first node of each table's log-uniform part (-1: none) with
       log x there and nodes per unit of log x; nguess + 1 interval indices of bg_chi at
       chi = k chi_max / nguess (stored as doubles) */
    int tab_f0[CE_INTG_NSLOT];
    double tab_lx0[CE_INTG_NSLOT], tab_lsc[CE_INTG_NSLOT];
    int tab_ref;   /* a slot whose nodes every present standard table shares, or -1 */
    int nguess;
    const double *chi_guess;
    /* workspace of (CE_INTG_VAL + CE_INTG_PAIR) (ntab_values + CE_INTG_NBGC nbg)
       doubles, 64-byte aligned, for the Akima value records and node pairs, and
       CE_INTG_ARGS_BYTES of device memory for the kernel's arguments */
    int ntab_values;
    double *coef;
    void *args_dev;
    int nren0;
    const double *ren0_x, *ren0_y;
    int nren;
    const double *ren_chi;
    const double *ren_rec;   /* 16-byte aligned, per node (x fastest): {value, d/dx, d/dy, d2/dxdy} */
    double Om, H0h;
} ce_intg_dev;
/* one workgroup per (z, mu, sep): out [nz][nmu][nsep][2] = (single, double); rule =
   {p1, n1, pu, nu, ps, ns}, d_gl = the n1-, nu- and ns-point Gauss-Legendre rules on
   [-1, 1] (nodes then weights, in that order) */
int ce_launch_integrated(void *stream, const ce_intg_dev *in, unsigned flags, int nz, const double *d_z, int nmu,
                         const double *d_mu, int nsep, const double *d_sep, const int *rule, const double *d_gl,
                         double *d_out);

#ifdef __cplusplus
}
#endif
#endif
