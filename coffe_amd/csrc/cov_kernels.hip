// 2D FFTLog covariance integrals (gfx950, FP64).
//
// The reference computes D_{l1 l2}(r1, r2) = int dk k^2 P^p j_l1(k r1) j_l2(k r2) with a
// 2D FFTLog of the DIAGONAL matrix diag(k^3 P^p / dlnk) (covariance.c:357-470,
// twobessel.c:16-113): 2D r2c, separable window and gamma-ratio kernel, conj, 2D c2r
// (FFTW semantics: complex inverse along the first axis, then c2r along the second
// with the imaginary parts of its DC and Nyquist columns dropped), then a bicubic
// spline evaluated at sep x sep.
//
// For a diagonal input d_t the 2D r2c is D[(m+n) mod N] with D the 1D DFT of d, and
// the whole transform collapses, row by row, to two 1D FFTs:
//     Y_i(n) = B[n] sum_t d_t Ahat((i+t) mod N) e^{+2 pi i n t/N},  n <= N/2,
//     row_i  = c2r(Y_i),
// with A[m] = W_row(m) conj(H1[m]), B[n] = W_col(n) conj(H2[n]) (window times the
// conjugated kernel) and Ahat(q) = sum_m A[m] e^{+2 pi i m q/N}, both spectrum-
// independent.  Only the rows the bicubic needs are computed, and the bicubic
// itself is applied as two sparse linear functionals: the Hermite weights of the
// query and the spline-derivative weights of the nodes (which depend only on the
// r grid) give value(m, n) = sum_i U_m(i) sum_j V_n(j) row_i(j).
//
// Kernels:
//   k_cov_ahat   Ahat_l for each multipole (one workgroup each)
//   k_cov_diag   d_t = k^3 P(k)^p / dlnk / k^nu1 / k^nu2 at the log grid (Akima P)
//   k_cov_rows   per (spectrum, power, pair, row): the two FFTs, the output scale
//                pi / (16 N^2 r_j^nu2 r_i^nu1) and the column functionals V_n
//   k_cov_cols   the row functionals U_m and the (2l1+1)(2l2+1) {2,1}/pi^2 scale
//   k_cov_assemble  Gaussian covariance of the multipoles (covariance.c:796-1060)
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "ce_device.h"
#include "ce_fft.h"
#include "ce_internal.h"

namespace {

using namespace ce;

__global__ void __launch_bounds__(1024) k_cov_ahat(int N, ce_radix_plan rp, const cd *__restrict__ tw,
                                                   const cd *__restrict__ A, cd *__restrict__ ahat)
{
    extern __shared__ cd smem[];
    const int T = blockDim.x, tid = threadIdx.x;
    cd *lt = smem + lds_data_len(N);
    load_twiddles(lt, N, as_global(tw), 1);
    const cd *a = A + (int64_t)blockIdx.x * N;
    for (int i = tid; i < N; i += T) smem[pidx(i)] = cconj(a[i]);
    lds_barrier();
    fft_lds10(smem, N, rp.code, lt);
    cd *o = ahat + (int64_t)blockIdx.x * N;
    for (int i = tid; i < N; i += T) o[i] = cconj(smem[pidx(i)]);
}

struct diag_args {
    int N;
    const int64_t *in_off;   // [nspec]
    const int32_t *n_in;     // [nspec]
    const double *k, *p;     // concatenated spectra
    const double *kg;        // log grid k_t
    const double *k3;        // k_t^3
    const double *kn1, *kn2; // k_t^nu1, k_t^nu2
    double step;
    double *d;               // [nspec][2][N]
    int *status;
};

__global__ void __launch_bounds__(256) k_cov_diag(diag_args a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (t >= a.N) return;
    const int64_t off = a.in_off[s];
    const double P = ce::akima_general(a.k + off, a.p + off, a.n_in[s], a.kg[t]);
    // fk = k^3 P^power / dlnk, then / k^nu1 / k^nu2 (covariance.c:391-403, twobessel.c:49-55)
    const double d1 = a.k3[t] * P / a.step / a.kn1[t] / a.kn2[t];
    const double d2 = a.k3[t] * (P * P) / a.step / a.kn1[t] / a.kn2[t];
    double *o = a.d + (int64_t)s * 2 * a.N;
    o[t] = d1;
    o[a.N + t] = d2;
    if (!isfinite(d1) || !isfinite(d2)) a.status[0] = 1;
}

struct rows_args {
    int N, M;
    ce_radix_plan rpN, rpM;
    const cd *tw;            // e^{-2 pi i k/N}, k < N
    const cd *ahat;          // [nl][N]
    const cd *B;             // [nl][M+1]
    const double *d;         // [nspec][2][N]
    const int *pl1, *pl2;    // [npair] multipole slots
    int npair, nrows;
    const int *rows;         // [nrows] row indices
    const double *prn1, *prn2;  // r_j^nu1, r_j^nu2
    int R;
    const int *flo, *fcnt, *foff;   // functional n: first node, node count, weight offset
    const double *fw;        // weights times pi / (16 N^2 r_j^nu2)
    double *W;               // [nblk][npair][nrows][R]
};

// one workgroup per (row, pair, block); a block is one diagonal d (spectrum * 2 + power)
__global__ void __launch_bounds__(1024) k_cov_rows(rows_args a)
{
    extern __shared__ cd smem[];
    const int T = blockDim.x, tid = threadIdx.x;
    const int ir = blockIdx.x, pr = blockIdx.y, sp2 = blockIdx.z;
    const int N = a.N, M = a.M;
    const int i = a.rows[ir];
    gcd *tw = as_global(a.tw);
    cd *ltN = smem + lds_data_len(N);
    cd *ltM = ltN + 64 + ((N + 63) >> 6);
    load_twiddles(ltN, N, tw, 1);
    load_twiddles(ltM, M, tw, 2);
    gdouble *dv = as_global(a.d + (int64_t)sp2 * N);
    gcd *ah = as_global(a.ahat + (int64_t)a.pl1[pr] * N);
    gcd *B = as_global(a.B + (int64_t)a.pl2[pr] * (M + 1));
    // conj(u_t), u_t = d_t Ahat((i + t) mod N): the forward FFT then gives conj(V);
    // four rounds of loads in flight at a time
    for (int t0 = 0; t0 * T < N; t0 += 4) {
        cd av[4];
        double dd[4];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const int t = min(tid + (t0 + h) * T, N - 1);
            int q = i + t;
            if (q >= N) q -= N;
            av[h] = ldg(ah + q);
            dd[h] = dv[t];
        }
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const int t = tid + (t0 + h) * T;
            if (t < N) smem[pidx(t)] = cscale(cconj(av[h]), dd[h]);
        }
    }
    lds_barrier();
    fft_lds10(smem, N, a.rpN.code, ltN);
    // Y = B V on n <= M with Im Y_0 = Im Y_M = 0 (dropped by the c2r), merged into the
    // half-length sequence (stored conjugated for the inverse transform)
    for (int it0 = 0; it0 * T <= M / 2; it0 += 2) {
        cd b1[2], b2[2], w1[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int kk = min(tid + (it0 + h) * T, M / 2);
            b1[h] = ldg(B + kk);
            b2[h] = ldg(B + (M - kk));
            w1[h] = ldg(tw + kk);
        }
        // each pair (k, M - k) is read and rewritten by its own thread
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k1 = tid + (it0 + h) * T, k2 = M - k1;
            if (k1 <= M / 2) {
                cd Y1 = cmul(b1[h], cconj(smem[pidx(k1)]));
                cd Y2 = cmul(b2[h], cconj(smem[pidx(k2)]));
                if (k1 == 0) { Y1.y = 0.0; Y2.y = 0.0; }
                // e^{-2 pi i (M - k)/N} = -conj(e^{-2 pi i k/N}) (N = 2M)
                smem[pidx(k1)] = cconj(c2r_merge(Y1, Y2, w1[h]));
                if (k2 != k1 && k2 < M) smem[pidx(k2)] = cconj(c2r_merge(Y2, Y1, mk(-w1[h].x, w1[h].y)));
            }
        }
    }
    lds_barrier();
    fft_lds10(smem, M, a.rpM.code, ltM);
    // column functionals of this row: W[n] = sum_j V_n(j) row(j), the weights carrying
    // pi / (16 N^2 r_j^nu2); four threads per functional, combined by lane shuffles
    const double den_i = a.prn1[i];
    gdouble_out *Wo = as_global_out(a.W + (((int64_t)sp2 * a.npair + pr) * a.nrows + ir) * a.R);
    gdouble *fw = as_global(a.fw);
    for (int e = tid; e < 4 * a.R; e += T) {
        const int n = e >> 2, sub = e & 3;
        const int j0 = a.flo[n], cnt = a.fcnt[n];
        gdouble *w = fw + a.foff[n];
        double acc = 0.0;
        for (int q = sub; q < cnt; q += 4) {
            const int j = j0 + q;
            const cd z = smem[pidx(j >> 1)];
            acc += w[q] * ((j & 1) ? -z.y : z.x);
        }
        acc += __shfl_xor(acc, 1);
        acc += __shfl_xor(acc, 2);
        if (sub == 0) Wo[n] = acc / den_i;
    }
}

struct cols_args {
    int npair, nrows, R;
    const int *upos, *ucnt, *uoff;   // row functional m: first position in the row list, count, weight offset
    const double *uw;
    const double *fac;               // [2][npair]
    const double *W;                 // [nspec][2][npair][nrows][R]
    double *out;                     // [nspec][2][npair][R][R], value (m, n) at R n + m
};

__global__ void __launch_bounds__(256) k_cov_cols(cols_args a)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int RR = a.R * a.R;
    if (e >= RR) return;
    const int m = e / a.R, n = e % a.R;
    const int pr = blockIdx.y, sp2 = blockIdx.z;
    const int64_t blk = (int64_t)sp2 * a.npair + pr;
    const double *W = a.W + blk * a.nrows * a.R;
    const int p0 = a.upos[m], cnt = a.ucnt[m];
    const double *w = a.uw + a.uoff[m];
    double acc = 0.0;
    for (int q = 0; q < cnt; ++q) acc += w[q] * W[(int64_t)(p0 + q) * a.R + n];
    a.out[blk * RR + (int64_t)a.R * n + m] = acc * a.fac[(sp2 & 1) * a.npair + pr];
}

__global__ void k_fill_int(int *p, int n, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

struct asm_args {
    int nz, nl, R, npair;
    const int *ls;             // [nl]
    const int *pair_of;        // [nl][nl] -> pair index of (min, max)
    const double *sep;         // [R]
    const double *cc, *cm;     // [nz][nl][nl]
    const double *D1, *pscale, *vol, *pix;   // [nz]; pscale = 1 / (density1 density2 volume)
    int pd_same, pd_swap;      // Poisson deltas d(p1,p3) d(p2,p4) and d(p1,p4) d(p2,p3); both 0: no Poisson term
    int window, mixed, cosmic;
    const double *integ;       // [nspec][2][npair][R][R]
    double *out;               // [nspec][nz][nl][R][nl][R]
};

#define CE_ASM_T 32            // tile edge of k_cov_assemble

// k_cov_assemble: one workgroup (32 x 8 threads) per 32 x 32 tile (m, n) of one block
// (l1 = ls[i], l2 = ls[j]) of one spectrum.  The tile's integrals of P and P^2 are
// read once, coalesced along whichever index is contiguous in the stored block (m for
// i <= j, n for the transposed blocks i > j, covariance.c:779-788), through LDS, and
// the tile of every z bin is written with the threads along n (the output's fastest
// index).  Cosmic, mixed and Poisson terms as covariance.c:866-1031.
__global__ void __launch_bounds__(256) k_cov_assemble(asm_args a)
{
    __shared__ double t1[CE_ASM_T][CE_ASM_T + 1], t2[CE_ASM_T][CE_ASM_T + 1];   // [m][n]
    const int ntile = (a.R + CE_ASM_T - 1) / CE_ASM_T;
    const int tm = blockIdx.x / ntile, tn = blockIdx.x % ntile;
    const int blk = blockIdx.y, s = blockIdx.z;
    const int i = blk / a.nl, j = blk % a.nl;
    const int m0 = tm * CE_ASM_T, n0 = tn * CE_ASM_T;
    const int tx = threadIdx.x % CE_ASM_T, ty = threadIdx.x / CE_ASM_T;
    const int64_t RR = (int64_t)a.R * a.R;
    const int pr = a.pair_of[i * a.nl + j];
    const double *b1 = a.integ + ((int64_t)s * 2 * a.npair + pr) * RR;
    const double *b2 = a.integ + ((int64_t)s * 2 * a.npair + a.npair + pr) * RR;
    for (int r = ty; r < CE_ASM_T; r += 8) {
        if (i <= j) {   // stored at R n + m: threads along m
            const int m = m0 + tx, n = n0 + r;
            if (m < a.R && n < a.R) {
                t1[tx][r] = b1[(int64_t)a.R * n + m];
                t2[tx][r] = b2[(int64_t)a.R * n + m];
            }
        } else {        // the transpose of block (j, i), stored at R m + n: threads along n
            const int m = m0 + r, n = n0 + tx;
            if (m < a.R && n < a.R) {
                t1[r][tx] = b1[(int64_t)a.R * m + n];
                t2[r][tx] = b2[(int64_t)a.R * m + n];
            }
        }
    }
    __syncthreads();
    const int l1 = a.ls[i], l2 = a.ls[j];
    const int sgn_l2 = (l2 % 2 == 0) ? 1 : -1;
    const int sgn_d = (((l2 - l1) / 2) % 2 == 0) ? 1 : -1;
    const int sgn_s = ((l1 + l2) % 2 == 0) ? 1 : -1;
    const int n = n0 + tx;
    if (n >= a.R) return;
    const int64_t per = (int64_t)a.nz * a.nl * a.R * a.nl * a.R;
    double *out = a.out + (int64_t)s * per;
    const double pdelta = (double)(a.pd_same + sgn_l2 * a.pd_swap);
    for (int r = ty; r < CE_ASM_T; r += 8) {
        const int m = m0 + r;
        if (m >= a.R) break;
        const double dpk = t1[r][tx], dpk2 = t2[r][tx];
        const bool diag = l1 == l2 && m == n && pdelta != 0.0;   // the Poisson term's support
        for (int z = 0; z < a.nz; ++z) {
            const int zl = (z * a.nl + i) * a.nl + j;
            const double D = a.D1[z], V = a.vol[z];
            double cov = 0.0;
            if (diag) {
                const double sm = a.sep[m];
                double g;
                if (a.window) {
                    const double lo = sm - a.pix[z] / 2., hi = sm + a.pix[z] / 2.;
                    g = 3.0 / (hi * hi * hi - lo * lo * lo);
                } else {
                    g = 1. / sm / sm / a.pix[z];
                }
                cov += (2 * l1 + 1) / 4. / 3.14159265358979323846 * pdelta * g * a.pscale[z];
            }
            if (a.mixed) cov += sgn_d * sgn_s * D * D * dpk * a.cm[zl] / V;
            if (a.cosmic) cov += sgn_d * D * D * D * D * dpk2 * a.cc[zl] / V;
            out[((((int64_t)z * a.nl + i) * a.R + m) * a.nl + j) * a.R + n] = cov;
        }
    }
}

int set_lds(const void *fn, size_t bytes)
{
    static std::mutex mu;
    static const void *fns[4] = {nullptr};
    static size_t done[4] = {0};
    std::lock_guard<std::mutex> g(mu);
    int slot = -1;
    for (int i = 0; i < 4; ++i) {
        if (fns[i] == fn) { slot = i; break; }
        if (!fns[i]) { fns[i] = fn; slot = i; break; }
    }
    if (slot >= 0 && done[slot] >= bytes) return 0;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess && slot >= 0) done[slot] = bytes;
    return e == hipSuccess ? 0 : (int)e;
}

// ---- adjoint: d<gint, integrals>/dP for every input sample ----------------------------
// The integrals are linear in the diagonal d (per spectrum and power), so the adjoint
// runs the chain backwards: the row functionals transposed (Wbar), per row the column
// functionals transposed, the transposed c2r (an r2c, interior doubled), times conj(B),
// the forward FFT_N and Re(conj(Ahat(i+t)) u_t) summed over rows into dbar; then dbar
// through d = k^3 P^p / dlnk / k^nu1 / k^nu2 and the Akima spline onto the samples.

// Wbar[blk][ir][n] = fac sum_m U_m(ir) gint[blk][m][n] over the row functionals whose
// row window [upos_m, upos_m + cnt_m) holds ir; they lie in the range [mlo, mhi] that
// binary searches over the windows' monotone envelopes give
struct colsT_args {
    int npair, nrows, R;
    const int *upos, *ucnt, *uoff;
    const int *elo, *ehi;  // min of upos over m' >= m, max of upos + ucnt over m' <= m
    const double *uw, *fac;
    const double *gint;    // [nspec][2][npair][R][R], (m, n) at R n + m
    double *Wbar;          // [nspec][2][npair][nrows][R]
};

#define CE_COLS_T_RT 32     // rows per workgroup
#define CE_COLS_T_MC 16     // functionals per LDS tile
// one workgroup per (32 rows, pair, block); thread = column n (chunks of 256): a tile of
// G[n][m0 .. m0 + 16) is staged in LDS (rows padded to 17 doubles), the weights U_m(row)
// are wave-uniform loads, the 32 row sums stay in registers; m ascends as in a
// per-element loop, so the sums are the same
__global__ void __launch_bounds__(256) k_cov_cols_T(colsT_args a)
{
    __shared__ double gL[256 * (CE_COLS_T_MC + 1)];
    const int tid = threadIdx.x;
    const int r0 = blockIdx.x * CE_COLS_T_RT;
    const int r1 = min(r0 + CE_COLS_T_RT, a.nrows) - 1;
    const int pr = blockIdx.y, sp2 = blockIdx.z;
    const int64_t blk = (int64_t)sp2 * a.npair + pr;
    const double *g = a.gint + blk * a.R * a.R;
    // functionals whose envelope reaches rows r0..r1
    int lo = 0, hi = a.R;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a.ehi[mid] > r0) hi = mid; else lo = mid + 1;
    }
    const int mlo = lo;
    lo = 0; hi = a.R;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a.elo[mid] <= r1) lo = mid + 1; else hi = mid;
    }
    const int mhi = lo - 1;
    const double fac = a.fac[(sp2 & 1) * a.npair + pr];
    for (int nc = 0; nc < a.R; nc += 256) {
        double acc[CE_COLS_T_RT];
#pragma unroll
        for (int r = 0; r < CE_COLS_T_RT; ++r) acc[r] = 0.0;
        for (int m0 = mlo; m0 <= mhi; m0 += CE_COLS_T_MC) {
            const int mcn = min(CE_COLS_T_MC, mhi - m0 + 1);
            __syncthreads();
            for (int e = tid; e < 256 * CE_COLS_T_MC; e += 256) {
                const int nn = e / CE_COLS_T_MC, mc = e % CE_COLS_T_MC;
                const int ng = nc + nn;
                gL[nn * (CE_COLS_T_MC + 1) + mc] = (ng < a.R && mc < mcn) ? g[(int64_t)a.R * ng + m0 + mc] : 0.0;
            }
            __syncthreads();
            for (int mc = 0; mc < mcn; ++mc) {
                const int m = m0 + mc;
                const int up = a.upos[m], cn = a.ucnt[m], off = a.uoff[m];
                const double gv = gL[tid * (CE_COLS_T_MC + 1) + mc];
#pragma unroll
                for (int r = 0; r < CE_COLS_T_RT; ++r) {
                    const int q = r0 + r - up;
                    if (q >= 0 && q < cn) acc[r] += a.uw[off + q] * gv;
                }
            }
        }
        const int n = nc + tid;
        if (n < a.R)
#pragma unroll
            for (int r = 0; r < CE_COLS_T_RT; ++r)
                if (r0 + r < a.nrows) a.Wbar[(blk * a.nrows + r0 + r) * a.R + n] = acc[r] * fac;
    }
}

struct rowsT_args {
    int N, M;
    ce_radix_plan rpN, rpM;
    const cd *tw, *ahat, *B;
    const int *pl1, *pl2;
    int npair, nrows, chunk;   // rows per workgroup
    const int *rows;
    const double *prn1;
    int R;
    const int *flo, *fcnt, *foff;
    const double *fw;          // the scaled column weights (as the forward)
    const double *Wbar;        // [nblk][npair][nrows][R]
    double *dbar;              // [nblk][N], accumulated
};

#define CE_ROWS_T_ACC 8        // N <= 8 x 1024
// one workgroup per (row chunk, pair, block): every row of the chunk goes through the
// transposed transform in LDS, its contribution to dbar stays in registers (eight t per
// thread) until the chunk ends
__global__ void __launch_bounds__(1024) k_cov_rows_T(rowsT_args a)
{
    extern __shared__ cd smem[];
    const int T = blockDim.x, tid = threadIdx.x;
    const int c = blockIdx.x, pr = blockIdx.y, sp2 = blockIdx.z;
    const int N = a.N, M = a.M;
    gcd *tw = as_global(a.tw);
    cd *ltN = smem + lds_data_len(N);
    cd *ltM = ltN + 64 + ((N + 63) >> 6);
    load_twiddles(ltN, N, tw, 1);
    load_twiddles(ltM, M, tw, 2);
    gcd *ah = as_global(a.ahat + (int64_t)a.pl1[pr] * N);
    gcd *B = as_global(a.B + (int64_t)a.pl2[pr] * (M + 1));
    gdouble *fw = as_global(a.fw);
    double acc[CE_ROWS_T_ACC];
#pragma unroll
    for (int h = 0; h < CE_ROWS_T_ACC; ++h) acc[h] = 0.0;
    const int r0 = c * a.chunk, r1 = min(r0 + a.chunk, a.nrows);
    for (int ir = r0; ir < r1; ++ir) {
        const int i = a.rows[ir];
        gdouble *Wb = as_global(a.Wbar + (((int64_t)sp2 * a.npair + pr) * a.nrows + ir) * a.R);
        const double den_i = a.prn1[i];
        lds_barrier();   // the previous row's reads of smem are done
        for (int q = tid; q < M + 1; q += T) smem[pidx(q)] = mk(0.0, 0.0);
        lds_barrier();
        // rbar(j) = sum_n Wbar_n fw_n(j) / den_i, four threads per functional
        for (int e = tid; e < 4 * a.R; e += T) {
            const int n = e >> 2, sub = e & 3;
            const int j0 = a.flo[n], cnt = a.fcnt[n];
            gdouble *w = fw + a.foff[n];
            const double wn = Wb[n] / den_i;
            for (int q = sub; q < cnt; q += 4) atomicAdd(&lds_real(smem, j0 + q), w[q] * wn);
        }
        lds_barrier();
        fft_lds10(smem, M, a.rpM.code, ltM);
        // Ybar = r2c(rbar) (interior doubled), Vbar = conj(B) Ybar on k <= M, 0 above
        for (int k1 = tid; k1 <= M / 2; k1 += T) {
            const int k2 = M - k1;
            const cd Z1 = smem[pidx(k1)];
            const cd Z2 = smem[pidx(k2 == M ? 0 : k2)];
            const cd w1 = ldg(tw + k1), w2 = mk(-w1.x, w1.y);   // e^{-2 pi i (M - k)/N} = -conj(e^{-2 pi i k/N})
            cd R1 = r2c_split(Z1, Z2, w1);
            cd R2 = r2c_split(Z2, Z1, w2);
            R1 = (k1 == 0) ? mk(R1.x, 0.0) : cscale(R1, 2.0);
            R2 = (k2 == M) ? mk(R2.x, 0.0) : cscale(R2, 2.0);
            const cd b1 = ldg(B + k1), b2 = ldg(B + k2);
            // each pair (k, M - k) is read and rewritten by its own thread
            smem[pidx(k1)] = cmul(cconj(b1), R1);
            if (k2 != k1) smem[pidx(k2)] = cmul(cconj(b2), R2);
        }
        for (int q = M + 1 + tid; q < N; q += T) smem[pidx(q)] = mk(0.0, 0.0);
        lds_barrier();
        fft_lds10(smem, N, a.rpN.code, ltN);
        // dbar_t += Re(conj(Ahat(i + t)) u_t)
#pragma unroll
        for (int h = 0; h < CE_ROWS_T_ACC; ++h) {
            const int t = tid + h * T;
            if (t < N) {
                int q = i + t;
                if (q >= N) q -= N;
                const cd av = ldg(ah + q), u = smem[pidx(t)];
                acc[h] += av.x * u.x + av.y * u.y;
            }
        }
    }
    double *db = a.dbar + (int64_t)sp2 * N;
#pragma unroll
    for (int h = 0; h < CE_ROWS_T_ACC; ++h) {
        const int t = tid + h * T;
        if (t < N && r1 > r0) atomicAdd(&db[t], acc[h]);
    }
}

// value and derivatives with respect to six samples (forward-mode, for the Akima VJP)
struct dual6 {
    double v, d[6];
};
__device__ __forceinline__ dual6 dconst(double x)
{
    dual6 r;
    r.v = x;
#pragma unroll
    for (int q = 0; q < 6; ++q) r.d[q] = 0.0;
    return r;
}
__device__ __forceinline__ dual6 dadd(const dual6 &a, const dual6 &b)
{
    dual6 r;
    r.v = a.v + b.v;
#pragma unroll
    for (int q = 0; q < 6; ++q) r.d[q] = a.d[q] + b.d[q];
    return r;
}
__device__ __forceinline__ dual6 dsub(const dual6 &a, const dual6 &b)
{
    dual6 r;
    r.v = a.v - b.v;
#pragma unroll
    for (int q = 0; q < 6; ++q) r.d[q] = a.d[q] - b.d[q];
    return r;
}
__device__ __forceinline__ dual6 dmul(const dual6 &a, const dual6 &b)
{
    dual6 r;
    r.v = a.v * b.v;
#pragma unroll
    for (int q = 0; q < 6; ++q) r.d[q] = a.d[q] * b.v + a.v * b.d[q];
    return r;
}
__device__ __forceinline__ dual6 dscale(const dual6 &a, double s)
{
    dual6 r;
    r.v = a.v * s;
#pragma unroll
    for (int q = 0; q < 6; ++q) r.d[q] = a.d[q] * s;
    return r;
}
__device__ __forceinline__ dual6 ddiv(const dual6 &a, const dual6 &b)
{
    dual6 r;
    r.v = a.v / b.v;
#pragma unroll
    for (int q = 0; q < 6; ++q) r.d[q] = (a.d[q] - r.v * b.d[q]) / b.v;
    return r;
}
__device__ __forceinline__ dual6 dabs(const dual6 &a) { return a.v < 0 ? dscale(a, -1.0) : a; }

// the Akima value at xq in interval j of (x, y) (ce::akima_at, GSL's non-rounded ends)
// with its derivatives with respect to y[base .. base + 5] (every sample it reads)
__device__ dual6 akima_dual(const double *__restrict__ x, const double *__restrict__ y, int n, int j, double xq,
                            int base)
{
    auto Y = [&](int idx) {
        dual6 r = dconst(y[idx]);
        const int q = idx - base;
        if (q >= 0 && q < 6) r.d[q] = 1.0;
        return r;
    };
    auto sl = [&](int a0) { return dscale(dsub(Y(a0 + 1), Y(a0)), 1.0 / (x[a0 + 1] - x[a0])); };
    auto slope = [&](int a0) -> dual6 {
        if (a0 >= 0 && a0 <= n - 2) return sl(a0);
        if (a0 < 0) {
            const dual6 m0 = sl(0), m1 = sl(1);
            return a0 == -1 ? dsub(dscale(m0, 2.0), m1) : dsub(dscale(m0, 3.0), dscale(m1, 2.0));
        }
        const dual6 mn2 = sl(n - 2), mn3 = sl(n - 3);
        return a0 == n - 1 ? dsub(dscale(mn2, 2.0), mn3) : dsub(dscale(mn2, 3.0), dscale(mn3, 2.0));
    };
    const dual6 mm2 = slope(j - 2), mm1 = slope(j - 1), m0 = slope(j), mp1 = slope(j + 1), mp2 = slope(j + 2);
    const dual6 ne = dadd(dabs(dsub(mp1, m0)), dabs(dsub(mm1, mm2)));
    const double delx = xq - x[j];
    if (ne.v == 0.0) return dadd(Y(j), dscale(m0, delx));
    const double h = x[j + 1] - x[j];
    const dual6 nen = dadd(dabs(dsub(mp2, mp1)), dabs(dsub(m0, mm1)));
    const dual6 alpha = ddiv(dabs(dsub(mm1, mm2)), ne);
    dual6 tn;
    if (nen.v == 0.0) {
        tn = m0;
    } else {
        const dual6 an = ddiv(dabs(dsub(m0, mm1)), nen);
        tn = dadd(dmul(dsub(dconst(1.0), an), m0), dmul(an, mp1));
    }
    const dual6 b = dadd(dmul(dsub(dconst(1.0), alpha), mm1), dmul(alpha, m0));
    const dual6 cc = dscale(dsub(dsub(dscale(m0, 3.0), dscale(b, 2.0)), tn), 1.0 / h);
    const dual6 dd = dscale(dsub(dadd(b, tn), dscale(m0, 2.0)), 1.0 / (h * h));
    // y_j + delx (b + delx (c + d delx))
    return dadd(Y(j), dscale(dadd(b, dscale(dadd(cc, dscale(dd, delx)), delx)), delx));
}

struct diagT_args {
    int N;
    const int64_t *in_off;
    const int32_t *n_in;
    const double *k, *p, *kg, *k3, *kn1, *kn2;
    double step;
    const double *dbar;   // [nspec][2][N]
    double *gpk;          // like p, accumulated
};

// one thread per grid point: Pbar = dbar_1 k^3/dlnk/k^nu1/k^nu2 + dbar_2 2P k^3/...,
// then the Akima VJP onto the six samples of its stencil
__global__ void __launch_bounds__(256) k_cov_diag_T(diagT_args a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (t >= a.N) return;
    const int64_t off = a.in_off[s];
    const int n = a.n_in[s];
    const double *x = a.k + off, *y = a.p + off;
    const double xq = a.kg[t];
    if (!(xq >= x[0] && xq <= x[n - 1])) return;
    const int j = ce::interval_bisect(x, 0, n - 1, xq);
    const int base = max(0, min(j - 2, n - 6));
    const dual6 P = akima_dual(x, y, n, j, xq, base);
    const double sc = a.k3[t] / a.step / a.kn1[t] / a.kn2[t];
    const double *db = a.dbar + (int64_t)s * 2 * a.N;
    const double pbar = db[t] * sc + db[a.N + t] * 2.0 * P.v * sc;
#pragma unroll
    for (int q = 0; q < 6; ++q)
        if (P.d[q] != 0.0) atomicAdd(&a.gpk[off + base + q], pbar * P.d[q]);
}

__global__ void k_fill_double(double *p, int64_t n, double v)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

extern "C" size_t ce_cov_lds_bytes(int N)
{
    return sizeof(cd) * ((size_t)lds_data_len(N) + 64 + (size_t)((N + 63) >> 6) + 64 + (size_t)((N / 2 + 63) >> 6));
}

extern "C" int ce_launch_cov_ahat(void *stream, int nl, int N, const ce_radix_plan *rp, const double *d_tw,
                                  const double *d_A, double *d_ahat)
{
    const size_t lds = sizeof(cd) * ((size_t)lds_data_len(N) + 64 + (size_t)((N + 63) >> 6));
    int st = set_lds((const void *)k_cov_ahat, lds);
    if (st) return st;
    hipLaunchKernelGGL(k_cov_ahat, dim3(nl), dim3(1024), lds, (hipStream_t)stream, N, *rp,
                       reinterpret_cast<const cd *>(d_tw), reinterpret_cast<const cd *>(d_A),
                       reinterpret_cast<cd *>(d_ahat));
    return (int)hipGetLastError();
}

extern "C" int ce_launch_cov_diag(void *stream, int nspec, int N, const int64_t *d_in_off, const int32_t *d_n_in,
                                  const double *d_k, const double *d_pk, const double *d_kg, const double *d_k3,
                                  const double *d_kn1, const double *d_kn2, double step, double *d_d, int *d_status)
{
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, (hipStream_t)stream, d_status, 1, 0);
    diag_args a;
    a.N = N; a.in_off = d_in_off; a.n_in = d_n_in; a.k = d_k; a.p = d_pk; a.kg = d_kg; a.k3 = d_k3;
    a.kn1 = d_kn1; a.kn2 = d_kn2; a.step = step; a.d = d_d; a.status = d_status;
    hipLaunchKernelGGL(k_cov_diag, dim3((N + 255) / 256, nspec), dim3(256), 0, (hipStream_t)stream, a);
    return (int)hipGetLastError();
}

extern "C" int ce_launch_cov_rows(void *stream, int nblk, int N, const ce_radix_plan *rpN, const ce_radix_plan *rpM,
                                  const double *d_tw, const double *d_ahat, const double *d_B, const double *d_d,
                                  int npair, const int *d_pl1, const int *d_pl2, int nrows, const int *d_rows,
                                  const double *d_prn1, const double *d_prn2, int R, const int *d_flo,
                                  const int *d_fcnt, const int *d_foff, const double *d_fw, double *d_W)
{
    const size_t lds = ce_cov_lds_bytes(N);
    int st = set_lds((const void *)k_cov_rows, lds);
    if (st) return st;
    rows_args a;
    a.N = N; a.M = N / 2; a.rpN = *rpN; a.rpM = *rpM;
    a.tw = reinterpret_cast<const cd *>(d_tw);
    a.ahat = reinterpret_cast<const cd *>(d_ahat);
    a.B = reinterpret_cast<const cd *>(d_B);
    a.d = d_d; a.pl1 = d_pl1; a.pl2 = d_pl2; a.npair = npair; a.nrows = nrows; a.rows = d_rows;
    a.prn1 = d_prn1; a.prn2 = d_prn2; a.R = R; a.flo = d_flo; a.fcnt = d_fcnt; a.foff = d_foff; a.fw = d_fw;
    a.W = d_W;
    hipLaunchKernelGGL(k_cov_rows, dim3(nrows, npair, nblk), dim3(1024), lds, (hipStream_t)stream, a);
    return (int)hipGetLastError();
}

extern "C" int ce_launch_cov_cols(void *stream, int nspec, int npair, int nrows, int R, const int *d_upos,
                                  const int *d_ucnt, const int *d_uoff, const double *d_uw, const double *d_fac,
                                  const double *d_W, double *d_out)
{
    cols_args a;
    a.npair = npair; a.nrows = nrows; a.R = R; a.upos = d_upos; a.ucnt = d_ucnt; a.uoff = d_uoff; a.uw = d_uw;
    a.fac = d_fac; a.W = d_W; a.out = d_out;
    const int RR = R * R;
    hipLaunchKernelGGL(k_cov_cols, dim3((RR + 255) / 256, npair, 2 * nspec), dim3(256), 0, (hipStream_t)stream, a);
    return (int)hipGetLastError();
}

extern "C" int ce_launch_cov_assemble(void *stream, int nspec, int nz, int nl, int R, int npair, const int *d_ls,
                                      const int *d_pair_of, const double *d_sep, const double *d_cc,
                                      const double *d_cm, const double *d_D1, const double *d_pscale,
                                      const double *d_vol, const double *d_pix, int pd_same, int pd_swap, int window,
                                      int mixed, int cosmic, const double *d_integ, double *d_out)
{
    asm_args a;
    memset(&a, 0, sizeof(a));
    a.nz = nz; a.nl = nl; a.R = R; a.npair = npair; a.ls = d_ls; a.pair_of = d_pair_of; a.sep = d_sep;
    a.cc = d_cc; a.cm = d_cm; a.D1 = d_D1; a.pscale = d_pscale; a.vol = d_vol; a.pix = d_pix;
    a.pd_same = pd_same; a.pd_swap = pd_swap; a.window = window;
    a.mixed = mixed; a.cosmic = cosmic; a.integ = d_integ; a.out = d_out;
    const int ntile = (R + CE_ASM_T - 1) / CE_ASM_T;
    hipLaunchKernelGGL(k_cov_assemble, dim3((unsigned)(ntile * ntile), (unsigned)(nl * nl), (unsigned)nspec),
                       dim3(256), 0, (hipStream_t)stream, a);
    return (int)hipGetLastError();
}

// ---- adjoint of the assembly ------------------------------------------------------
// The covariance is linear in the stored integrals and in the per-(z, l1, l2)
// coefficients cm (mixed) and cc (cosmic).  k_cov_assemble_T: one thread per stored
// integral element (spectrum, pair (a, b), n_s, m_s) sums the cotangent over the z bins
// and the blocks that read it, (a, b) at (m_s, n_s) and, for a < b, (b, a) at (n_s, m_s).
__global__ void __launch_bounds__(256) k_cov_assemble_T(asm_args a, const double *__restrict__ gout,
                                                        double *__restrict__ ginteg)
{
    const int RR = a.R * a.R;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)a.npair * RR) return;
    const int s = blockIdx.y;
    const int pr = (int)(e / RR);
    const int idx = (int)(e % RR);
    const int ns = idx / a.R, ms = idx % a.R;
    int ia = 0, ib = 0;
    for (int i = 0; i < a.nl; ++i)
        for (int j = i; j < a.nl; ++j)
            if (a.pair_of[i * a.nl + j] == pr) { ia = i; ib = j; }
    const int64_t per = (int64_t)a.nz * a.nl * a.R * a.nl * a.R;
    const double *g = gout + (int64_t)s * per;
    double g1 = 0.0, g2 = 0.0;
    for (int z = 0; z < a.nz; ++z) {
        const double D = a.D1[z], V = a.vol[z];
        for (int o = 0; o < ((ia < ib) ? 2 : 1); ++o) {
            const int i = o ? ib : ia, j = o ? ia : ib;
            const int m = o ? ns : ms, n = o ? ms : ns;
            const int l1 = a.ls[i], l2 = a.ls[j];
            const int sgn_d = (((l2 - l1) / 2) % 2 == 0) ? 1 : -1;
            const int sgn_s = ((l1 + l2) % 2 == 0) ? 1 : -1;
            const int zl = (z * a.nl + i) * a.nl + j;
            const double go = g[((((int64_t)z * a.nl + i) * a.R + m) * a.nl + j) * a.R + n];
            if (a.mixed) g1 += sgn_d * sgn_s * D * D * a.cm[zl] / V * go;
            if (a.cosmic) g2 += sgn_d * D * D * D * D * a.cc[zl] / V * go;
        }
    }
    double *gi = ginteg + (int64_t)s * 2 * a.npair * RR;
    gi[(int64_t)pr * RR + idx] = g1;
    gi[((int64_t)a.npair + pr) * RR + idx] = g2;
}

// k_cov_coef_T: one workgroup per (spectrum, z, l1, l2) block; gcoef[..][0] = d<gout,
// cov>/d cm and gcoef[..][1] = d<gout, cov>/d cc of the block (tree sum in LDS)
__global__ void __launch_bounds__(256) k_cov_coef_T(asm_args a, const double *__restrict__ gout,
                                                    double *__restrict__ gcoef)
{
    __shared__ double red[2][256];
    const int blk = blockIdx.x, s = blockIdx.y, tid = threadIdx.x;
    const int j = blk % a.nl, i = (blk / a.nl) % a.nl, z = blk / (a.nl * a.nl);
    const int l1 = a.ls[i], l2 = a.ls[j];
    const int RR = a.R * a.R;
    const int pr = a.pair_of[i * a.nl + j];
    const double *base = a.integ + (int64_t)s * 2 * a.npair * RR;
    const int64_t per = (int64_t)a.nz * a.nl * a.R * a.nl * a.R;
    const double *g = gout + (int64_t)s * per;
    double t1 = 0.0, t2 = 0.0;
    for (int e = tid; e < RR; e += blockDim.x) {
        const int m = e / a.R, n = e % a.R;
        const int64_t idx = (i <= j) ? (int64_t)a.R * n + m : (int64_t)a.R * m + n;
        const double go = g[((((int64_t)z * a.nl + i) * a.R + m) * a.nl + j) * a.R + n];
        t1 += go * base[(int64_t)pr * RR + idx];
        t2 += go * base[((int64_t)a.npair + pr) * RR + idx];
    }
    red[0][tid] = t1;
    red[1][tid] = t2;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (tid < w) {
            red[0][tid] += red[0][tid + w];
            red[1][tid] += red[1][tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) {
        const double D = a.D1[z], V = a.vol[z];
        const int sgn_d = (((l2 - l1) / 2) % 2 == 0) ? 1 : -1;
        const int sgn_s = ((l1 + l2) % 2 == 0) ? 1 : -1;
        double *o = gcoef + (((int64_t)s * a.nz + z) * a.nl * a.nl + i * a.nl + j) * 2;
        o[0] = a.mixed ? sgn_d * sgn_s * D * D / V * red[0][0] : 0.0;
        o[1] = a.cosmic ? sgn_d * D * D * D * D / V * red[1][0] : 0.0;
    }
}

extern "C" int ce_launch_cov_assemble_vjp(void *stream, int nspec, int nz, int nl, int R, int npair, const int *d_ls,
                                          const int *d_pair_of, const double *d_cc, const double *d_cm,
                                          const double *d_D1, const double *d_vol, int mixed, int cosmic,
                                          const double *d_integ, const double *d_gout, double *d_ginteg,
                                          double *d_gcoef)
{
    asm_args a;
    memset(&a, 0, sizeof(a));
    a.nz = nz; a.nl = nl; a.R = R; a.npair = npair; a.ls = d_ls; a.pair_of = d_pair_of;
    a.cc = d_cc; a.cm = d_cm; a.D1 = d_D1; a.vol = d_vol;
    a.mixed = mixed; a.cosmic = cosmic; a.integ = d_integ;
    const int64_t nel = (int64_t)npair * R * R;
    hipLaunchKernelGGL(k_cov_assemble_T, dim3((unsigned)((nel + 255) / 256), nspec), dim3(256), 0,
                       (hipStream_t)stream, a, d_gout, d_ginteg);
    hipLaunchKernelGGL(k_cov_coef_T, dim3((unsigned)(nz * nl * nl), nspec), dim3(256), 0, (hipStream_t)stream, a,
                       d_gout, d_gcoef);
    return (int)hipGetLastError();
}

extern "C" int ce_launch_cov_vjp(void *stream, int nspec, int N, const ce_radix_plan *rpN, const ce_radix_plan *rpM,
                                 const double *d_tw, const double *d_ahat, const double *d_B, int npair,
                                 const int *d_pl1, const int *d_pl2, int nrows, const int *d_rows,
                                 const double *d_prn1, int R, const int *d_flo, const int *d_fcnt, const int *d_foff,
                                 const double *d_fws, const int *d_upos, const int *d_uelo, const int *d_uehi,
                                 const double *d_fw, const double *d_fac,
                                 const double *d_gint, double *d_Wbar, double *d_dbar, const int64_t *d_in_off,
                                 const int32_t *d_n_in, const double *d_k, const double *d_pk, const double *d_kg,
                                 const double *d_k3, const double *d_kn1, const double *d_kn2, double step,
                                 double *d_gpk, int64_t n_total)
{
    if (N > CE_ROWS_T_ACC * 1024) return (int)hipErrorInvalidValue;
    hipStream_t s = (hipStream_t)stream;
    colsT_args c;
    c.npair = npair; c.nrows = nrows; c.R = R; c.upos = d_upos; c.ucnt = d_fcnt; c.uoff = d_foff; c.uw = d_fw;
    c.elo = d_uelo; c.ehi = d_uehi;
    c.fac = d_fac; c.gint = d_gint; c.Wbar = d_Wbar;
    hipLaunchKernelGGL(k_cov_cols_T, dim3((nrows + CE_COLS_T_RT - 1) / CE_COLS_T_RT, npair, 2 * nspec), dim3(256), 0,
                       s, c);
    const int64_t nd = (int64_t)2 * nspec * N;
    hipLaunchKernelGGL(k_fill_double, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, s, d_dbar, nd, 0.0);
    hipLaunchKernelGGL(k_fill_double, dim3((unsigned)((n_total + 255) / 256)), dim3(256), 0, s, d_gpk, n_total, 0.0);
    const size_t lds = ce_cov_lds_bytes(N);
    int st = set_lds((const void *)k_cov_rows_T, lds);
    if (st) return st;
    rowsT_args a;
    a.N = N; a.M = N / 2; a.rpN = *rpN; a.rpM = *rpM;
    a.tw = reinterpret_cast<const cd *>(d_tw);
    a.ahat = reinterpret_cast<const cd *>(d_ahat);
    a.B = reinterpret_cast<const cd *>(d_B);
    a.pl1 = d_pl1; a.pl2 = d_pl2; a.npair = npair; a.nrows = nrows; a.rows = d_rows; a.prn1 = d_prn1;
    a.R = R; a.flo = d_flo; a.fcnt = d_fcnt; a.foff = d_foff; a.fw = d_fws; a.Wbar = d_Wbar; a.dbar = d_dbar;
    // about 2 x 256 workgroups in all: rows per workgroup
    const int want = (512 + npair * 2 * nspec - 1) / (npair * 2 * nspec);
    const int nchunk = want < 1 ? 1 : (want > nrows ? nrows : want);
    a.chunk = (nrows + nchunk - 1) / nchunk;
    const int grid_x = (nrows + a.chunk - 1) / a.chunk;
    hipLaunchKernelGGL(k_cov_rows_T, dim3(grid_x, npair, 2 * nspec), dim3(1024), lds, s, a);
    diagT_args d;
    d.N = N; d.in_off = d_in_off; d.n_in = d_n_in; d.k = d_k; d.p = d_pk; d.kg = d_kg; d.k3 = d_k3;
    d.kn1 = d_kn1; d.kn2 = d_kn2; d.step = step; d.dbar = d_dbar; d.gpk = d_gpk;
    hipLaunchKernelGGL(k_cov_diag_T, dim3((N + 255) / 256, nspec), dim3(256), 0, s, d);
    return (int)hipGetLastError();
}
