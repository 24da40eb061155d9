// twofast_1bessel_noalloc: the drop-in 1D FFTLog (twofast_1bessel_status, the
// reference's twofast.c:362-483) for callers that cannot allocate device memory.
//
// The existing entry point keeps its spectrum-independent grid tables in a cache on
// the engine and its per-call workspace in a scratch buffer that it grows with
// hipMalloc.  This variant allocates nothing: inputs and outputs are the caller's
// device arrays, and everything else -- the grid tables (recomputed on every call),
// the resampled spectrum, the r2c spectrum and the kernels' descriptors -- lives in
// the file-scope device array `scratch` below, of a fixed size (FFT length N <= 16384,
// input length <= 65536; about 1.2 MB per GPU, held until the process ends).  Host
// code reaches it only through the address hipGetSymbolAddress returns for the
// current device, fetched once per device on the first call there.
//
// The kernels are the existing entry point's (k_grid_tables, k_resample, k_prepare,
// k_tables in full-table mode); only the twiddle factors and the log grid
// (kmax/kmin)^(i/N) are computed on the device here instead of on the host, so the
// results agree with twofast_1bessel_status to rounding.
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <mutex>

#include "coffe_engine.h"
#include "ce_host.h"
#include "ce_internal.h"

namespace {

constexpr int kMaxN = 16384;              // largest FFT length
constexpr size_t kMaxIn = 65536;          // largest input length
constexpr size_t kMetaBytes = 16384;      // descriptors, chunks and the status word
constexpr size_t kMaxChunks = (kMaxIn - 1 + CE_CHUNK - 1) / CE_CHUNK;
// grid tables p (N), twiddles (2N), bias*window (N), U (2(M+1)), output scale (N);
// per call the resampled spectrum (2N) and the r2c spectrum (2(M+1)): 9N + 4 doubles
constexpr size_t kScratchDoubles = 9 * (size_t)kMaxN + 4 + kMetaBytes / sizeof(double);
constexpr int kMaxDevices = 64;

constexpr size_t align16(size_t b) { return (b + 15) & ~(size_t)15; }
constexpr size_t kMetaUsed = align16(sizeof(ce_spec_dev)) + align16(sizeof(ce_grid_dev)) +
                             align16(sizeof(ce_table_desc)) + align16(sizeof(ce_unit_dev)) + 16 +
                             kMaxChunks * align16(sizeof(ce_chunk_dev));
static_assert(kMetaUsed <= kMetaBytes, "descriptor region too small");

}  // namespace

// the scratch of twofast_1bessel_noalloc: one copy per GPU, shared by every call on it
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

std::mutex g_lock;                        // serialises the calls of this process
double *g_scratch[kMaxDevices] = {nullptr};

struct setup_args {
    int N, nchunk;
    int64_t n;
    double ratio;
    double *p, *tw;
    const double *kin, *pin;
    int *status;
    ce_spec_dev spec;
    ce_grid_dev grid;
    ce_table_desc tab;
    ce_unit_dev unit;
    ce_spec_dev *d_spec;
    ce_grid_dev *d_grid;
    ce_table_desc *d_tab;
    ce_unit_dev *d_unit;
    ce_chunk_dev *d_chunks;
};

__global__ void k_noalloc_clear(int *status) { *status = 0; }

// the log grid and twiddles of length N, the descriptors, the chunks of the input
// intervals, and the finiteness of the input samples (status bit 1)
__global__ void k_noalloc_setup(setup_args a)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        *a.d_spec = a.spec;
        *a.d_grid = a.grid;
        *a.d_tab = a.tab;
        *a.d_unit = a.unit;
    }
    if (i < a.N) {
        a.p[i] = pow(a.ratio, (double)i / a.N);
        double s, c;
        sincospi(-2.0 * (double)i / a.N, &s, &c);
        a.tw[2 * i] = c;
        a.tw[2 * i + 1] = s;
    }
    if (i < a.nchunk) {
        const int nint = (int)a.n - 1;
        ce_chunk_dev c;
        c.spec = 0;
        c.grid = 0;
        c.in_off = 0;
        c.ws_off = 0;
        c.n_in = (int32_t)a.n;
        c.id = (int32_t)i;
        c.a = (int32_t)i * CE_CHUNK;
        c.b = (int32_t)((i + 1) * CE_CHUNK < nint ? (i + 1) * CE_CHUNK : nint);
        a.d_chunks[i] = c;
    }
    if (i < a.n && (!isfinite(a.kin[i]) || !isfinite(a.pin[i]))) atomicOr(a.status, 1);
}

// a non-finite gamma-ratio kernel (status bit 2; the reference exits, twofast.c:210)
__global__ void k_noalloc_check(const double *u, int n, int *status)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !isfinite(u[i])) atomicOr(status, 2);
}

__global__ void k_noalloc_outx(const double *p, int N, double r0, double *out_x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) out_x[i] = r0 * p[i];
}

// the device address of `scratch` on the current device, looked up once per device
int scratch_address(double **out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return COFFE_ENODEVICE;
    if (dev < 0 || dev >= kMaxDevices) return COFFE_EUNSUPPORTED;
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return COFFE_EDEVICE;
        g_scratch[dev] = static_cast<double *>(ptr);
    }
    *out = g_scratch[dev];
    return COFFE_OK;
}

}  // namespace

extern "C" int twofast_1bessel_noalloc(double *d_output_x, double *d_output_y, size_t output_len,
                                       const double *d_input_x, const double *d_input_y, size_t input_len, double l,
                                       double nu, double r0, double k0, double kmin, double kmax, void *stream)
{
    if (!d_output_x || !d_output_y || !d_input_x || !d_input_y || input_len < 5) return COFFE_EINVAL;
    if (input_len > kMaxIn) return COFFE_EUNSUPPORTED;
    if (output_len < 4 || (output_len & 1) || output_len > (size_t)kMaxN) return COFFE_EUNSUPPORTED;
    if (!(kmin > 0) || !(kmax > kmin) || !(k0 > 0) || !(r0 > 0) || !isfinite(kmax)) return COFFE_EINVAL;
    const int N = (int)output_len, M = N / 2;
    ce_radix_plan rp;
    if (ce_radix_schedule(M, &rp) != COFFE_OK) return COFFE_EUNSUPPORTED;
    // window limits of twofast_fft_input (twofast.c:283-289)
    const double xmax = k0 * pow(kmax / kmin, (double)(N - 1) / N);
    const double xl = exp(0.46) * kmin, xr = exp(-0.46) * xmax;
    if (!(kmin <= xl && xl <= xr && xr <= xmax)) return COFFE_EDOMAIN;

    std::lock_guard<std::mutex> guard(g_lock);
    double *base = nullptr;
    int st = scratch_address(&base);
    if (st != COFFE_OK) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);

    // scratch layout (doubles): p | tw | bw | U | os | ws | xf | descriptors
    double *p = base, *tw = base + N, *bw = base + 3 * (size_t)N, *U = base + 4 * (size_t)N;
    double *os = U + 2 * (size_t)(M + 1), *ws = os + N, *xf = ws + 2 * (size_t)N;
    char *meta = reinterpret_cast<char *>(xf + 2 * (size_t)(M + 1));
    ce_spec_dev *d_spec = reinterpret_cast<ce_spec_dev *>(meta);
    ce_grid_dev *d_grid = reinterpret_cast<ce_grid_dev *>(meta + align16(sizeof(ce_spec_dev)));
    ce_table_desc *d_tab =
        reinterpret_cast<ce_table_desc *>(reinterpret_cast<char *>(d_grid) + align16(sizeof(ce_grid_dev)));
    ce_unit_dev *d_unit =
        reinterpret_cast<ce_unit_dev *>(reinterpret_cast<char *>(d_tab) + align16(sizeof(ce_table_desc)));
    int *d_status = reinterpret_cast<int *>(reinterpret_cast<char *>(d_unit) + align16(sizeof(ce_unit_dev)));
    ce_chunk_dev *d_chunks = reinterpret_cast<ce_chunk_dev *>(reinterpret_cast<char *>(d_status) + 16);
    const int nchunk = (int)((input_len - 1 + CE_CHUNK - 1) / CE_CHUNK);

    // one table (l, nu), q + nu its single q' class, no r^(n-l) factor
    ce_table_desc tab;
    tab.n = nu;
    tab.l = l;
    tab.qidx = 0;
    tab.rpow = 0;
    const double qnu = ce_selectqnu(l, nu);
    const double qprime = qnu + nu;

    setup_args a;
    a.N = N;
    a.nchunk = nchunk;
    a.n = (int64_t)input_len;
    a.ratio = kmax / kmin;
    a.p = p;
    a.tw = tw;
    a.kin = d_input_x;
    a.pin = d_input_y;
    a.status = d_status;
    memset(&a.spec, 0, sizeof(a.spec));
    a.spec.n_in = (int32_t)input_len;
    memset(&a.grid, 0, sizeof(a.grid));
    a.grid.rp = rp;
    a.grid.N = N;
    a.grid.M = M;
    a.grid.kmin = kmin;
    a.grid.kmax = kmax;
    a.grid.G = log(kmax / kmin);
    a.grid.k0 = k0;
    a.grid.r0 = r0;
    a.grid.p = p;
    a.grid.tw = tw;
    a.grid.bw[0] = bw;
    a.grid.U[0] = U;
    a.grid.os[0] = os;
    a.tab = tab;
    a.unit.spec = 0;
    a.unit.tab = 0;
    a.d_spec = d_spec;
    a.d_grid = d_grid;
    a.d_tab = d_tab;
    a.d_unit = d_unit;
    a.d_chunks = d_chunks;

    int threads = 512;
    while (threads < M / 8) threads *= 2;
    const int64_t nset = (int64_t)N > (int64_t)input_len ? (int64_t)N : (int64_t)input_len;
    hipLaunchKernelGGL(k_noalloc_clear, dim3(1), dim3(1), 0, s, d_status);
    hipLaunchKernelGGL(k_noalloc_setup, dim3((unsigned)((nset + 255) / 256)), dim3(256), 0, s, a);
    if (hipGetLastError() != hipSuccess) return COFFE_EDEVICE;
    double *Ut[1] = {U}, *ost[1] = {os}, *bwt[1] = {bw};
    if (ce_launch_grid_tables(s, N, kmin, kmax, k0, r0, p, 1, &tab, &qnu, 1, &qprime, Ut, ost, bwt))
        return COFFE_EDEVICE;
    hipLaunchKernelGGL(k_noalloc_check, dim3((2 * (M + 1) + 255) / 256), dim3(256), 0, s, U, 2 * (M + 1), d_status);
    // non-finite input samples or kernel values stop the call before the spline kernels
    // run on them (the existing entry point checks the samples on the host first)
    int h_status = 0;
    if (hipMemcpyAsync(&h_status, d_status, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) return COFFE_EDEVICE;
    if (hipStreamSynchronize(s) != hipSuccess) return COFFE_EDEVICE;
    if (h_status) return COFFE_ENONFINITE;
    if (ce_launch_resample(s, nchunk, d_chunks, d_spec, d_grid, d_input_x, d_input_y, ws, NULL)) return COFFE_EDEVICE;
    if (ce_launch_prepare(s, 1, d_spec, d_grid, 1, ws, xf, threads, M)) return COFFE_EDEVICE;
    if (ce_launch_tables(s, 1, d_unit, d_spec, d_grid, d_tab, 1, xf, 0, NULL, NULL, NULL, d_output_y, N, NULL, NULL, 1,
                         1, NULL, NULL, threads, M, NULL, 0, 0))
        return COFFE_EDEVICE;
    hipLaunchKernelGGL(k_noalloc_outx, dim3((N + 255) / 256), dim3(256), 0, s, p, N, r0, d_output_x);
    if (hipGetLastError() != hipSuccess) return COFFE_EDEVICE;
    return hipStreamSynchronize(s) == hipSuccess ? COFFE_OK : COFFE_EDEVICE;
}
