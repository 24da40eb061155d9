// Complex FFTs of any length on device memory, for the transform lengths the fused
// LDS kernels do not take (more than 8192 complex points, or a prime factor above 5),
// and the FFTLog drop-in built on them (twofast.c:362-483 for those lengths).
//
//   direct      L = 2^a 3^b 5^c <= 8192: one LDS Stockham transform per line
//   four-step   L = L1 L2, both direct: L2 column transforms of length L1 (stride L2)
//               with the twiddles e^{-2 pi i n2 k1 / L} folded into their store, L1 row
//               transforms of length L2, then the transposed result gathered into
//               natural order
//   Bluestein   any other L: the chirp w_n = e^{-i pi n^2 / L}, a length-P cyclic
//               convolution (P >= 2L - 1, direct or four-step) against the transform
//               of the conjugate chirp, made once per plan
//
// Transforms are forward (e^{-2 pi i k n / L}) and unnormalised; the inverse is the
// conjugate of the forward transform of the conjugate.  Twiddles of the line
// transforms come from correctly rounded host tables (ce_twiddles); the four-step and
// chirp factors from sincospi of an exactly reduced integer argument.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>

#include "ce_fft.h"
#include "ce_host.h"
#include "ce_internal.h"

namespace {

using namespace ce;

constexpr int LINE_LEN_MAX = 8192;   // the longest LDS line transform

struct lines_args {
    const cd *in;
    cd *out;
    int L;
    uint32_t code;
    const cd *tw;           // e^{-2 pi i k / L}, k < L
    int64_t line_stride;    // between lines (same for in and out)
    int64_t elem_stride;    // between the elements of a line
    int64_t P;              // four-step twiddle: element k of line j times e^{-2 pi i j k / P} (0: none)
    int64_t batch_stride;   // between the transforms of a batch (blockIdx.y)
};

// e^{-2 pi i m / P} for 0 <= m < P
__device__ __forceinline__ cd unit_root(int64_t m, int64_t P)
{
    double s, c;
    sincospi(-2.0 * (double)m / (double)P, &s, &c);
    return mk(c, s);
}

// one line transform per workgroup of at least L/8 threads (the LDS passes' bound on
// butterflies per thread); lines are disjoint, so in == out is allowed
__global__ void __launch_bounds__(1024) k_lines(lines_args a)
{
    extern __shared__ cd smem[];
    const int L = a.L, tid = threadIdx.x, T = blockDim.x;
    const int64_t j = blockIdx.x, bofs = (int64_t)blockIdx.y * a.batch_stride;
    cd *lt = smem + lds_data_len(L);
    load_twiddles(lt, L, as_global(a.tw), 1);
    gcd *in = as_global(a.in + bofs + j * a.line_stride);
    for (int i = tid; i < L; i += T) smem[pidx(i)] = ldg(in + i * a.elem_stride);
    lds_barrier();
    fft_lds(smem, L, a.code, lt, tid, T);
    lds_barrier();
    gcd_out *out = as_global_out(a.out + bofs + j * a.line_stride);
    for (int k = tid; k < L; k += T) {
        cd v = smem[pidx(k)];
        if (a.P) v = cmul(v, unit_root((j * k) % a.P, a.P));
        stg(out + k * a.elem_stride, v);
    }
}

// natural order from the four-step's transposed result: out[k1 + L1 k2] = in[k1 L2 + k2],
// for nb transforms of L = L1 L2 points one after another
__global__ void k_untranspose(const cd *__restrict__ in, cd *__restrict__ out, int L1, int L2, int64_t nb)
{
    const int64_t L = (int64_t)L1 * L2;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nb * L; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = e / L, k = e - b * L;
        const int64_t k1 = k % L1, k2 = k / L1;
        stg(as_global_out(out + e), ldg(as_global(in + b * L + k1 * L2 + k2)));
    }
}

// e^{-i pi n^2 / L} with n^2 reduced modulo 2L exactly
__device__ __forceinline__ cd chirp(int64_t n, int64_t L)
{
    const int64_t m = (n % (2 * L)) * (n % (2 * L)) % (2 * L);
    double s, c;
    sincospi(-(double)m / (double)L, &s, &c);
    return mk(c, s);
}

// Bluestein input: a_n = x_n w_n (n < L), 0 up to P, for nb transforms
__global__ void k_blue_in(const cd *__restrict__ x, cd *__restrict__ a, int64_t L, int64_t P, int64_t nb)
{
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nb * P; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = e / P, n = e - b * P;
        stg(as_global_out(a + e), n < L ? cmul(ldg(as_global(x + b * L + n)), chirp(n, L)) : mk(0.0, 0.0));
    }
}

// the conjugate chirp b_n = conj(w_n) at n and P - n (n < L), 0 elsewhere
__global__ void k_blue_kernel(cd *__restrict__ b, int64_t L, int64_t P)
{
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < P; n += (int64_t)gridDim.x * blockDim.x) {
        cd v = mk(0.0, 0.0);
        if (n < L) v = cconj(chirp(n, L));
        else if (n > P - L) v = cconj(chirp(P - n, L));
        stg(as_global_out(b + n), v);
    }
}

// the product with the kernel's transform, conjugated for the inverse transform
__global__ void k_blue_mul(cd *__restrict__ a, const cd *__restrict__ bhat, int64_t P, int64_t nb)
{
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nb * P; e += (int64_t)gridDim.x * blockDim.x)
        stg(as_global_out(a + e), cconj(cmul(ldg(as_global(a + e)), ldg(as_global(bhat + e % P)))));
}

// X_k = w_k conj(c_k) / P
__global__ void k_blue_out(const cd *__restrict__ c, cd *__restrict__ x, int64_t L, int64_t P, int64_t nb)
{
    const double inv = 1.0 / (double)P;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nb * L; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = e / L, k = e - b * L;
        stg(as_global_out(x + e), cmul(chirp(k, L), cscale(cconj(ldg(as_global(c + b * P + k))), inv)));
    }
}

// ---- the FFTLog drop-in on these transforms (twofast.c:362-483) -------------------

// c_i = (bw_i S_i, 0): the windowed, biased input on the log grid (twofast.c:275-290)
__global__ void k_tf_in(const double *__restrict__ S, const double *__restrict__ bw, cd *__restrict__ c, int N)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        stg(as_global_out(c + i), mk(bw[i] * S[i], 0.0));
}

// from Y = FFT_N(c): X_m = conj(Y_m) U_m for m <= N/2 (imaginary parts of X_0 and
// X_{N/2} dropped, as a real c2r input), then D = conj of the Hermitian extension of
// X, in place: D_m = conj(X_m), D_{N-m} = X_m (twofast.c:426-446)
__global__ void k_tf_mid(cd *__restrict__ y, const cd *__restrict__ U, int N)
{
    const int M = N / 2;
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m <= M; m += gridDim.x * blockDim.x) {
        cd X = cmul(cconj(ldg(as_global(y + m))), ldg(as_global(U + m)));
        if (m == 0 || m == M) X.y = 0.0;
        stg(as_global_out(y + m), cconj(X));
        if (m > 0 && m < M) stg(as_global_out(y + (N - m)), X);
    }
}

// out_i = os_i Re(E_i), E = FFT_N(D): the c2r output times the output prefactor
// (twofast.c:415-421, 466-477)
__global__ void k_tf_out(const cd *__restrict__ e, const double *__restrict__ os, double *__restrict__ out, int N)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        out[i] = os[i] * e[i].x;
}

// ---- the 2D FFTLog rows of the covariance on these transforms ---------------------
// (the general-length counterpart of the fused row kernel: the same row sums, transform
// for transform).  Row e of a chunk is e = (blk npair + pair) nrows + ir.

// work[c][t] = conj(Ahat_l1((i + t) mod N)) d_blk[t], i = rows[ir]
__global__ void k_covg_in(const cd *__restrict__ ahat, const double *__restrict__ d, const int *__restrict__ pl1,
                          const int *__restrict__ rows, int npair, int nrows, int N, int64_t e0, int64_t ne,
                          cd *__restrict__ work)
{
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < ne * N;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = idx / N, e = e0 + c;
        const int t = (int)(idx - c * N);
        const int ir = (int)(e % nrows), pr = (int)((e / nrows) % npair);
        const int64_t blk = e / ((int64_t)nrows * npair);
        int q = rows[ir] + t;
        if (q >= N) q -= N;
        const cd a = ldg(as_global(ahat + (int64_t)pl1[pr] * N + q));
        stg(as_global_out(work + idx), cscale(cconj(a), d[blk * N + t]));
    }
}

// Y_n = B_l2[n] conj(Z_n) for n <= N/2 (imaginary parts at 0 and N/2 dropped, as a c2r
// input), then the conjugate of its Hermitian extension, in place (D_n = conj Y_n,
// D_{N-n} = Y_n)
__global__ void k_covg_mid(cd *__restrict__ work, const cd *__restrict__ B, const int *__restrict__ pl2, int npair,
                           int nrows, int N, int64_t e0, int64_t ne)
{
    const int M = N / 2;
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < ne * (M + 1);
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = idx / (M + 1);
        const int m = (int)(idx - c * (M + 1));
        const int pr = (int)(((e0 + c) / nrows) % npair);
        cd *w = work + c * N;
        cd Y = cmul(ldg(as_global(B + (int64_t)pl2[pr] * (M + 1) + m)), cconj(ldg(as_global(w + m))));
        if (m == 0 || m == M) Y.y = 0.0;
        stg(as_global_out(w + m), cconj(Y));
        if (m > 0 && m < M) stg(as_global_out(w + (N - m)), Y);
    }
}

// W[e][n] = sum_q fw[foff[n] + q] row(flo[n] + q) / prn1[i], row = Re of the second
// transform (the weights carry the output scale pi / (16 N^2 r_j^nu2))
__global__ void k_covg_func(const cd *__restrict__ work, const int *__restrict__ rows, const double *__restrict__ prn1,
                            int nrows, int R, const int *__restrict__ flo, const int *__restrict__ fcnt,
                            const int *__restrict__ foff, const double *__restrict__ fw, int N, int64_t e0,
                            int64_t ne, double *__restrict__ W)
{
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < ne * R;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = idx / R, e = e0 + c;
        const int n = (int)(idx - c * R);
        const cd *row = work + c * N;
        const int j0 = flo[n], cnt = fcnt[n];
        const double *w = fw + foff[n];
        double acc = 0.0;
        for (int q = 0; q < cnt; ++q) acc += w[q] * row[j0 + q].x;
        W[e * R + n] = acc / prn1[rows[(int)(e % nrows)]];
    }
}

__global__ void k_conj(const cd *__restrict__ in, cd *__restrict__ out, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        stg(as_global_out(out + i), cconj(ldg(as_global(in + i))));
}

int blocks_for(int64_t n)
{
    const int64_t b = (n + 255) / 256;
    return (int)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

}  // namespace

// ---- plans -----------------------------------------------------------------------

struct ce_anyfft {
    using cd = ce::cd;
    int64_t L;
    int kind;   // 0 direct, 1 four-step, 2 Bluestein
    int L1, L2;
    ce_radix_plan rp1, rp2;
    cd *tw1, *tw2;            // line twiddle tables (device)
    cd *tmp;                  // four-step gather buffer (device), tmp_cap transforms
    int64_t tmp_cap;
    struct ce_anyfft *sub;    // Bluestein: the length-P transform
    int64_t P;
    cd *bhat;                 // Bluestein: FFT_P(b) (device)
    cd *work;                 // Bluestein: the length-P work arrays (device), work_cap transforms
    int64_t work_cap;
};

namespace {

bool line_ok(int64_t L, ce_radix_plan *rp)
{
    return L >= 1 && L <= LINE_LEN_MAX && (L == 1 || ce_radix_schedule((int)L, rp) == COFFE_OK);
}

bool smooth(int64_t n)
{
    for (int p : {2, 3, 5})
        while (n % p == 0) n /= p;
    return n == 1;
}

// a split L = L1 L2 with both factors line lengths, L1 the largest such divisor up to sqrt(L)
bool split(int64_t L, int *L1, int *L2)
{
    if (!smooth(L)) return false;
    for (int64_t a = (int64_t)std::sqrt((double)L) + 1; a >= 1; --a) {
        if (L % a) continue;
        const int64_t b = L / a;
        ce_radix_plan r1, r2;
        if (a <= LINE_LEN_MAX && b <= LINE_LEN_MAX && line_ok(a, &r1) && line_ok(b, &r2)) {
            *L1 = (int)a;
            *L2 = (int)b;
            return true;
        }
    }
    return false;
}

int twiddle_table(int L, cd **out)
{
    double *h = (double *)malloc(sizeof(double) * 2 * (size_t)L);
    if (!h) return COFFE_ENOMEM;
    ce_twiddles(L, h);
    int st = ce_dev_alloc((void **)out, sizeof(cd) * (size_t)L);
    if (st == COFFE_OK) st = ce_hip_ok(hipMemcpy(*out, h, sizeof(cd) * (size_t)L, hipMemcpyHostToDevice));
    free(h);
    return st;
}

// count lines of each of nb transforms (batch_stride apart)
int launch_lines(hipStream_t s, cd *x, int L, const ce_radix_plan &rp, const cd *tw, int64_t count,
                 int64_t line_stride, int64_t elem_stride, int64_t P, int64_t nb, int64_t batch_stride)
{
    if (count <= 0 || nb <= 0 || L == 1) return COFFE_OK;   // a length-1 transform is the identity
    if (nb > 65535) return COFFE_EUNSUPPORTED;
    lines_args a;
    a.in = x;
    a.out = x;
    a.L = L;
    a.code = rp.code;
    a.tw = tw;
    a.line_stride = line_stride;
    a.elem_stride = elem_stride;
    a.P = P;
    a.batch_stride = batch_stride;
    const size_t lds = ((size_t)lds_data_len(L) + 64 + (size_t)(L + 63) / 64) * sizeof(cd);
    static bool attr_set = false;
    if (!attr_set) {
        if (hipFuncSetAttribute((const void *)k_lines, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) !=
            hipSuccess)
            return COFFE_EDEVICE;
        attr_set = true;
    }
    int threads = 64;
    while (threads < L / 8) threads *= 2;
    hipLaunchKernelGGL(k_lines, dim3((unsigned)count, (unsigned)nb), dim3(threads), lds, s, a);
    return hipGetLastError() == hipSuccess ? COFFE_OK : COFFE_EDEVICE;
}

// a device buffer of at least n transforms of len points (the stream drained before an
// old buffer is released)
int reserve(cd **buf, int64_t *cap, int64_t n, int64_t len, hipStream_t s)
{
    if (*cap >= n) return COFFE_OK;
    if (*buf) {
        int st = ce_hip_ok(hipStreamSynchronize(s));
        if (st != COFFE_OK) return st;
        (void)hipFree(*buf);
        *buf = NULL;
        *cap = 0;
    }
    int st = ce_dev_alloc((void **)buf, sizeof(cd) * (size_t)(n * len));
    if (st == COFFE_OK) *cap = n;
    return st;
}

}  // namespace

extern "C" {

void ce_anyfft_destroy(ce_anyfft *p)
{
    if (!p) return;
    (void)hipFree(p->tw1);
    (void)hipFree(p->tw2);
    (void)hipFree(p->tmp);
    (void)hipFree(p->bhat);
    (void)hipFree(p->work);
    ce_anyfft_destroy(p->sub);
    free(p);
}

int ce_anyfft_forward_batch(ce_anyfft *p, void *d_x, int64_t nb, void *stream);

int ce_anyfft_create(int64_t L, ce_anyfft **out)
{
    *out = NULL;
    if (L < 1 || L > ((int64_t)1 << 26)) return COFFE_EUNSUPPORTED;
    ce_anyfft *p = (ce_anyfft *)calloc(1, sizeof(*p));
    if (!p) return COFFE_ENOMEM;
    p->L = L;
    int st = COFFE_OK;
    if (line_ok(L, &p->rp1)) {
        p->kind = 0;
        p->L1 = (int)L;
        if (L > 1) st = twiddle_table((int)L, &p->tw1);
    } else if (split(L, &p->L1, &p->L2)) {
        p->kind = 1;
        line_ok(p->L1, &p->rp1);
        line_ok(p->L2, &p->rp2);
        if (p->L1 > 1) st = twiddle_table(p->L1, &p->tw1);
        if (st == COFFE_OK && p->L2 > 1) st = twiddle_table(p->L2, &p->tw2);
    } else {
        p->kind = 2;
        int64_t P = 2 * L - 1;
        int a, b;
        while (!(line_ok(P, &p->rp1) || split(P, &a, &b))) ++P;
        p->P = P;
        st = ce_anyfft_create(P, &p->sub);
        if (st == COFFE_OK) st = ce_dev_alloc((void **)&p->bhat, sizeof(cd) * (size_t)P);
        if (st == COFFE_OK) {
            hipLaunchKernelGGL(k_blue_kernel, dim3(blocks_for(P)), dim3(256), 0, 0, p->bhat, L, P);
            st = hipGetLastError() == hipSuccess ? COFFE_OK : COFFE_EDEVICE;
        }
        if (st == COFFE_OK) st = ce_anyfft_forward_batch(p->sub, p->bhat, 1, NULL);
        if (st == COFFE_OK) st = ce_hip_ok(hipDeviceSynchronize());
    }
    if (st != COFFE_OK) {
        ce_anyfft_destroy(p);
        return st;
    }
    *out = p;
    return COFFE_OK;
}

// nb transforms of L points one after another at d_x, in place
int ce_anyfft_forward_batch(ce_anyfft *p, void *d_x, int64_t nb, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    cd *x = (cd *)d_x;
    const int64_t L = p->L;
    if (nb <= 0) return COFFE_OK;
    if (p->kind == 0) return launch_lines(s, x, (int)L, p->rp1, p->tw1, 1, 0, 1, 0, nb, L);
    if (p->kind == 1) {
        const int L1 = p->L1, L2 = p->L2;
        // columns: line n2 = elements n1 L2 + n2, twiddled by e^{-2 pi i n2 k1 / L}
        int st = launch_lines(s, x, L1, p->rp1, p->tw1, L2, 1, L2, L, nb, L);
        // rows: line k1 = elements k1 L2 + k2
        if (st == COFFE_OK) st = launch_lines(s, x, L2, p->rp2, p->tw2, L1, L2, 1, 0, nb, L);
        if (st == COFFE_OK) st = reserve(&p->tmp, &p->tmp_cap, nb, L, s);
        if (st == COFFE_OK) {
            hipLaunchKernelGGL(k_untranspose, dim3(blocks_for(nb * L)), dim3(256), 0, s, x, p->tmp, L1, L2, nb);
            st = hipGetLastError() == hipSuccess ? COFFE_OK : COFFE_EDEVICE;
        }
        if (st == COFFE_OK)
            st = ce_hip_ok(hipMemcpyAsync(x, p->tmp, sizeof(cd) * (size_t)(nb * L), hipMemcpyDeviceToDevice, s));
        return st;
    }
    const int64_t P = p->P;
    int st = reserve(&p->work, &p->work_cap, nb, P, s);
    if (st != COFFE_OK) return st;
    hipLaunchKernelGGL(k_blue_in, dim3(blocks_for(nb * P)), dim3(256), 0, s, x, p->work, L, P, nb);
    if (hipGetLastError() != hipSuccess) return COFFE_EDEVICE;
    st = ce_anyfft_forward_batch(p->sub, p->work, nb, stream);
    if (st != COFFE_OK) return st;
    hipLaunchKernelGGL(k_blue_mul, dim3(blocks_for(nb * P)), dim3(256), 0, s, p->work, p->bhat, P, nb);
    if (hipGetLastError() != hipSuccess) return COFFE_EDEVICE;
    st = ce_anyfft_forward_batch(p->sub, p->work, nb, stream);
    if (st != COFFE_OK) return st;
    hipLaunchKernelGGL(k_blue_out, dim3(blocks_for(nb * L)), dim3(256), 0, s, p->work, x, L, P, nb);
    return hipGetLastError() == hipSuccess ? COFFE_OK : COFFE_EDEVICE;
}

int ce_anyfft_forward(ce_anyfft *p, void *d_x, void *stream) { return ce_anyfft_forward_batch(p, d_x, 1, stream); }

int ce_anyfft_kind(const ce_anyfft *p) { return p->kind; }
int64_t ce_anyfft_padded(const ce_anyfft *p) { return p->kind == 2 ? p->P : p->L; }

// the FFTLog stages around the two transforms; d_c is a length-N complex work array
int ce_launch_tf_in(void *stream, const double *d_S, const double *d_bw, void *d_c, int N)
{
    hipLaunchKernelGGL(k_tf_in, dim3(blocks_for(N)), dim3(256), 0, (hipStream_t)stream, d_S, d_bw, (cd *)d_c, N);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ce_launch_tf_mid(void *stream, void *d_c, const double *d_U, int N)
{
    hipLaunchKernelGGL(k_tf_mid, dim3(blocks_for(N / 2 + 1)), dim3(256), 0, (hipStream_t)stream, (cd *)d_c,
                       (const cd *)d_U, N);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ce_launch_tf_out(void *stream, const void *d_c, const double *d_os, double *d_out, int N)
{
    hipLaunchKernelGGL(k_tf_out, dim3(blocks_for(N)), dim3(256), 0, (hipStream_t)stream, (const cd *)d_c, d_os,
                       d_out, N);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// Ahat_l = conj(FFT_N(conj(A_l))) for nl arrays (the general-length k_cov_ahat)
int ce_general_cov_ahat(ce_anyfft *fft, void *stream, int nl, int N, const double *d_A, double *d_ahat)
{
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = (int64_t)nl * N;
    hipLaunchKernelGGL(k_conj, dim3(blocks_for(n)), dim3(256), 0, s, (const cd *)d_A, (cd *)d_ahat, n);
    if (hipGetLastError() != hipSuccess) return COFFE_EDEVICE;
    int st = ce_anyfft_forward_batch(fft, d_ahat, nl, stream);
    if (st != COFFE_OK) return st;
    hipLaunchKernelGGL(k_conj, dim3(blocks_for(n)), dim3(256), 0, s, (const cd *)d_ahat, (cd *)d_ahat, n);
    return hipGetLastError() == hipSuccess ? COFFE_OK : COFFE_EDEVICE;
}

// the general-length ce_launch_cov_rows: rows in chunks of at most 2^28 bytes of work,
// each chunk two batched length-N transforms; synchronises the stream
int ce_general_cov_rows(ce_anyfft *fft, void *stream, int nblk, int N, const double *d_ahat, const double *d_B,
                        const double *d_d, int npair, const int *d_pl1, const int *d_pl2, int nrows,
                        const int *d_rows, const double *d_prn1, int R, const int *d_flo, const int *d_fcnt,
                        const int *d_foff, const double *d_fw, double *d_W)
{
    hipStream_t s = (hipStream_t)stream;
    const int64_t total = (int64_t)nblk * npair * nrows;
    int64_t chunk = ((int64_t)1 << 28) / ((int64_t)sizeof(cd) * N);
    if (chunk < 1) chunk = 1;
    if (chunk > 65535) chunk = 65535;
    if (chunk > total) chunk = total;
    cd *work = NULL;
    int st = ce_dev_alloc((void **)&work, sizeof(cd) * (size_t)(chunk * N));
    for (int64_t e0 = 0; e0 < total && st == COFFE_OK; e0 += chunk) {
        const int64_t ne = total - e0 < chunk ? total - e0 : chunk;
        hipLaunchKernelGGL(k_covg_in, dim3(blocks_for(ne * N)), dim3(256), 0, s, (const cd *)d_ahat, d_d, d_pl1,
                           d_rows, npair, nrows, N, e0, ne, work);
        st = hipGetLastError() == hipSuccess ? COFFE_OK : COFFE_EDEVICE;
        if (st == COFFE_OK) st = ce_anyfft_forward_batch(fft, work, ne, stream);
        if (st == COFFE_OK) {
            hipLaunchKernelGGL(k_covg_mid, dim3(blocks_for(ne * (N / 2 + 1))), dim3(256), 0, s, work,
                               (const cd *)d_B, d_pl2, npair, nrows, N, e0, ne);
            st = hipGetLastError() == hipSuccess ? COFFE_OK : COFFE_EDEVICE;
        }
        if (st == COFFE_OK) st = ce_anyfft_forward_batch(fft, work, ne, stream);
        if (st == COFFE_OK) {
            hipLaunchKernelGGL(k_covg_func, dim3(blocks_for(ne * R)), dim3(256), 0, s, work, d_rows, d_prn1, nrows, R,
                               d_flo, d_fcnt, d_foff, d_fw, N, e0, ne, d_W);
            st = hipGetLastError() == hipSuccess ? COFFE_OK : COFFE_EDEVICE;
        }
    }
    const int sync = ce_hip_ok(hipStreamSynchronize(s));
    if (st == COFFE_OK) st = sync;
    (void)hipFree(work);
    return st;
}

}  // extern "C"
