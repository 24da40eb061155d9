// HIP kernels of the FFTLog engine (gfx950, FP64).
//
// A workgroup (or, in k_tables, each of its thread groups) owns one FFT at a time.
// The length-M complex transform (M = N/2, N the real FFTLog length) lives in LDS as
// interleaved doubles and is computed by self-sorting Stockham passes (ce_fft.h): a
// single LDS buffer of 16*M bytes (M <= 8192 -> 128 KiB of the 160 KiB per CU).  Real
// transforms use the half-length complex transform plus the split step.  The spline
// kernels run one wavefront per chunk of input intervals.
//
// Kernels (reference steps they implement):
//   k_grid_tables  spectrum-independent tables of one (N, kmin, kmax) grid:
//                  bias*window (twofast.c:275-290), gamma-ratio kernel u_l with the
//                  Fourier window and 1/L folded in (twofast.c:196-224, 345-353,
//                  426-446), output prefactor times r^(n-l) (twofast.c:415-421,
//                  integrals.c:514-522).
//   k_resample     per chunk of input intervals: natural cubic spline of P(k)
//                  (twofast.c:399-401) and its values on the log grid (275-290).
//   k_prepare      per spectrum: bias * window * S, r2c, conj (twofast.c:252-360).
//   k_tables       per (spectrum, table) unit, units of one length two to a workgroup:
//                  kernel multiply, c2r (twofast.c:426-477),
//                  then either the full table or Akima at the separations
//                  (common.c:704-710 on the table of integrals.c:564-567); the
//                  adjoint mode adds the Akima VJP and the transposed c2r.
//   k_project      multipoles = sum_t A[z,l,t,r] I_t(r) (the mu-integrated weights).
//   k_adj_c2r      per spectrum: transposed r2c and bias -> cotangent of S(x_i);
//                  in the composite adjoint call also the projection of k_project.
//   k_spline_adj   per chunk: transposed resample and spline -> dP.
#include <hip/hip_runtime.h>

#include <mutex>

#include "ce_fft.h"
#include "ce_internal.h"

namespace {

using namespace ce;

// ---- complex helpers for the gamma-ratio kernel -------------------------------
__device__ __forceinline__ cd c_log(cd z) { return mk(log(hypot(z.x, z.y)), atan2(z.y, z.x)); }
__device__ __forceinline__ cd c_exp(cd z)
{
    double s, c;
    sincos(z.y, &s, &c);
    const double e = exp(z.x);
    return mk(e * c, e * s);
}
__device__ __forceinline__ cd c_div(cd a, cd b)
{
    const double d = b.x * b.x + b.y * b.y;
    return mk((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__device__ __forceinline__ cd c_sin(cd z)
{
    double s, c;
    sincos(z.x, &s, &c);
    return mk(s * cosh(z.y), c * sinh(z.y));
}

// complex log Gamma, Lanczos g = 7 with 8 coefficients (twofast.c:97-144)
__device__ cd lgamma_lanczos7(cd z)
{
    const double PI = 3.14159265358979323846;
    const double coef[8] = {676.5203681218851,    -1259.1392167224028,  771.32342877765313,
                            -176.61502916214059,  12.507343278686905,   -0.13857109526572012,
                            9.9843695780195716e-6, 1.5056327351493116e-7};
    bool reflect = z.x < 0.5;
    cd zz = reflect ? mk(1.0 - z.x, -z.y) : z;
    cd x = mk(0.99999999999980993, 0.0);
    zz.x -= 1.0;
    for (int i = 0; i < 8; ++i) x = cadd(x, c_div(mk(coef[i], 0.0), mk(zz.x + (double)i + 1.0, zz.y)));
    cd t = mk(zz.x + 7.5, zz.y);
    cd lt = c_log(t);
    cd r = cadd(mk(log(2 * PI) / 2.0, 0.0), cmul(mk(zz.x + 0.5, zz.y), lt));
    r = cadd(csub(r, t), c_log(x));
    if (reflect) {
        cd sn = c_sin(mk(PI * z.x, PI * z.y));
        r = csub(csub(mk(log(PI), 0.0), c_log(sn)), r);
    }
    return r;
}

// twofast_mql: 2^(n-1) sqrt(pi) exp(lgG((1+l+n)/2) - lgG((2+l-n)/2)) alpha^(i t - q), n = q - 1 - i t
__device__ cd mql(double t, double q, double l, double alpha)
{
    const double SQRT_PI = 1.7724538509055160272981674833411451827975494561224;
    cd n = mk(q - 1.0, -t);
    // cpow(2, n - 1) = cexp((n - 1) clog 2)
    const double l2 = log(2.0);
    cd p2 = c_exp(mk((n.x - 1.0) * l2, n.y * l2));
    cd g1 = lgamma_lanczos7(mk((1.0 + l + n.x) / 2.0, n.y / 2.0));
    cd g2 = lgamma_lanczos7(mk((2.0 + l - n.x) / 2.0, -n.y / 2.0));
    cd unl = cscale(cmul(p2, c_exp(csub(g1, g2))), SQRT_PI);
    const double la = log(alpha);
    cd pa = c_exp(mk(-q * la, t * la));
    return cmul(pa, unl);
}

// twofast_window (twofast.c:54-90); the limits are validated on the host
__device__ __forceinline__ double kwindow(double v, double xmin, double xmax, double xl, double xr)
{
    const double PI = 3.14159265358979323846;
    if (v > xl && v < xr && v > xmin && v < xmax) return 1.0;
    if (v <= xmin || v >= xmax) return 0.0;
    double r = (v <= xl) ? (v - xmin) / (xl - xmin) : (xmax - v) / (xmax - xr);
    return r - sin(2 * PI * r) / 2. / PI;
}

struct grid_tab_args {
    int N;
    double kmin, kmax, k0, r0;
    const double *p;
    int ntab, nq;
    int qidx[CE_MAX_TABLES], rpow[CE_MAX_TABLES];
    double n[CE_MAX_TABLES], l[CE_MAX_TABLES];
    double qnu[CE_MAX_TABLES], qprime[2];
    double *U[CE_MAX_TABLES], *os[CE_MAX_TABLES], *bw[2];
};

__global__ void k_grid_tables(grid_tab_args a)
{
    const double PI = 3.14159265358979323846;
    const int N = a.N, M = N / 2;
    const double ratio = a.kmax / a.kmin;
    const double G = log(ratio);
    const double L = 2 * PI * N / G;
    const double k0 = a.k0;
    const double xmax = k0 * a.p[N - 1];
    const double xl = exp(0.46) * a.kmin, xr = exp(-0.46) * xmax;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) {
        const double x = k0 * a.p[i];
        const double w = kwindow(x, a.kmin, xmax, xl, xr);
        for (int q = 0; q < a.nq; ++q) a.bw[q][i] = pow(ratio, (3. - a.qprime[q]) * i / N) * w;
        for (int t = 0; t < a.ntab; ++t) {
            const double qp = a.qprime[a.qidx[t]];
            const double nu = a.n[t];
            double v = k0 * k0 * k0 * pow(ratio, -(qp * i / N)) / PI / pow(a.r0 * k0, nu) / G;
            if (a.rpow[t] && a.n[t] >= a.l[t]) v *= pow(a.r0 * a.p[i], a.n[t] - a.l[t]);
            a.os[t][i] = v;
        }
    }
    if (i <= M) {
        const double wf = kwindow(k0 * a.p[M - 1 + i], a.kmin, xmax, xl, xr) / L;
        for (int t = 0; t < a.ntab; ++t) {
            cd u = mql(2 * PI * i / G, a.qnu[t], a.l[t], k0 * a.r0);
            a.U[t][2 * i] = u.x * wf;
            a.U[t][2 * i + 1] = u.y * wf;
        }
    }
}

// ---- natural cubic spline and resampling, chunked over the input nodes ---------
// A chunk owns the input intervals [a, b) of one spectrum and is handled by one
// wavefront.  It stages the nodes [a - W - 1, b + W + 1] in LDS, solves the natural-
// spline system (unknowns c_i = y''_i / 2, i = 1..n-2, twofast.c:399-401) restricted
// to the nodes [max(1, a - W), min(n - 2, b + W)], and keeps the coefficients of
// [a, b].  Cutting the system W rows away perturbs the kept rows by at most
// prod gamma <= 2^-W relative (the matrix is strictly diagonally dominant, every
// LDL^T multiplier gamma is below 1/2), 5e-20 for W = 64; the global ends carry the
// exact natural boundary c_0 = c_{n-1} = 0.

#define CE_SPLINE_HALO 64

// 1/x within about one ulp without the division sequence: the hardware estimate
// refined by two Newton steps (no VCC, so independent reciprocals overlap); for
// normal x of moderate size, as the spacings and diagonals here are
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// the barrier of a one-wavefront workgroup: a wavefront's LDS operations complete
// in issue order, so only the compiler has to keep the accesses on their side (and
// global loads still in flight stay in flight)
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// rows per lane of the one-wavefront solve: 64 lanes hold the largest truncated
// system, CE_CHUNK + 2 CE_SPLINE_HALO + 1 rows
#define CE_SOLVE_M 11
static_assert(64 * CE_SOLVE_M >= CE_CHUNK + 2 * CE_SPLINE_HALO + 1, "spline solve capacity");
static_assert(CE_CHUNK_NODES >= CE_CHUNK + 2 * CE_SPLINE_HALO + 3, "spline staging capacity");

// The natural-spline system of the staged nodes xs (row r = node u0 + r, staged at
// r + 1; h_r = xs[r+1] - xs[r]):
//     h_r c_{r-1} + 2 (h_r + h_{r+1}) c_r + h_{r+1} c_{r+1} = d_r,   r < ns,
// without the left term in row 0 and the right term in row ns - 1, solved by one
// wavefront as a partitioned system.  Lane p owns rows [p M, p M + M); the last is
// its separator s_p.  Its interior rows are eliminated in registers (Thomas, with
// the two coupling columns as extra right-hand sides):
//     x_j = y_j - v_j s_{p-1} - w_j s_p,
// so the separator rows form a 64-row tridiagonal system, solved exactly by six
// steps of cyclic reduction across the lanes (lane shuffles; each step forms
// bL bR row_i - a bR row_L - c bL row_R and rescales it by a power of two, which
// needs no division and is exact).  Rows past ns are identity rows.  rhs(j, r, hm,
// hp) gives d_r for the lane's row j (a constant after unrolling), r = lane M + j < ns,
// called in increasing j.  The solution goes to X[r], r < ns; every lane of the
// wavefront must call.
template <class Rhs>
__device__ __forceinline__ void wave_spline_solve(const double *xs, int ns, Rhs rhs, double *X)
{
    constexpr int M = CE_SOLVE_M;
    const int lane = threadIdx.x & 63;
    const int r0 = lane * M;
    double cp[M - 1], y[M - 1], v[M - 1], w[M - 1];
    double cprev = 0.0, yprev = 0.0, vprev = 0.0;
#pragma unroll
    for (int j = 0; j < M - 1; ++j) {
        const int r = r0 + j;
        double a = 0.0, b = 1.0, c = 0.0, d = 0.0;
        if (r < ns) {
            const double hm = xs[r + 1] - xs[r], hp = xs[r + 2] - xs[r + 1];
            a = (r > 0) ? hm : 0.0;
            c = (r < ns - 1) ? hp : 0.0;
            b = 2.0 * (hm + hp);
            d = rhs(j, r, hm, hp);
        }
        // row 0 of the lane: its left term multiplies s_{p-1} (right-hand side v)
        const double rd = rcp_nr(b - a * cprev);
        cp[j] = c * rd;
        y[j] = (d - a * yprev) * rd;
        v[j] = (j == 0 ? a : -a * vprev) * rd;
        cprev = cp[j];
        yprev = y[j];
        vprev = v[j];
    }
    // the right term of the last interior row multiplies s_p (right-hand side w)
    w[M - 2] = cp[M - 2];
#pragma unroll
    for (int j = M - 3; j >= 0; --j) {
        y[j] -= cp[j] * y[j + 1];
        v[j] -= cp[j] * v[j + 1];
        w[j] = -cp[j] * w[j + 1];
    }
    // separator row: a x_{pM+M-2} + b s_p + c x_{(p+1)M} = d with both neighbours
    // written through the separators
    double as = 0.0, bs = 1.0, cs = 0.0, ds = 0.0;
    {
        const int r = r0 + M - 1;
        if (r < ns) {
            const double hm = xs[r + 1] - xs[r], hp = xs[r + 2] - xs[r + 1];
            as = hm;
            cs = (r < ns - 1) ? hp : 0.0;
            bs = 2.0 * (hm + hp);
            ds = rhs(M - 1, r, hm, hp);
        }
    }
    double yn = __shfl_down(y[0], 1), vn = __shfl_down(v[0], 1), wn = __shfl_down(w[0], 1);
    if (lane == 63) { yn = 0.0; vn = 0.0; wn = 0.0; }
    double al = -as * v[M - 2];
    double be = bs - as * w[M - 2] - cs * vn;
    double ga = -cs * wn;
    double de = ds - as * y[M - 2] - cs * yn;
#pragma unroll
    for (int st = 1; st < 64; st <<= 1) {
        double aL = __shfl_up(al, st), bL = __shfl_up(be, st), cL = __shfl_up(ga, st), dL = __shfl_up(de, st);
        double aR = __shfl_down(al, st), bR = __shfl_down(be, st), cR = __shfl_down(ga, st), dR = __shfl_down(de, st);
        if (lane < st) { aL = 0.0; bL = 1.0; cL = 0.0; dL = 0.0; }
        if (lane + st >= 64) { aR = 0.0; bR = 1.0; cR = 0.0; dR = 0.0; }
        const double u = al * bR, q = ga * bL;
        const double ra = -u * aL;
        const double rb = be * bL * bR - u * cL - q * aR;
        const double rc = -q * cR;
        const double rdd = de * bL * bR - u * dL - q * dR;
        const int e = -__builtin_amdgcn_frexp_exp(rb);
        al = ldexp(ra, e);
        be = ldexp(rb, e);
        ga = ldexp(rc, e);
        de = ldexp(rdd, e);
    }
    const double sp = de / be;
    double sm = __shfl_up(sp, 1);
    if (lane == 0) sm = 0.0;
#pragma unroll
    for (int j = 0; j < M - 1; ++j)
        if (r0 + j < ns) X[r0 + j] = y[j] - v[j] * sm - w[j] * sp;
    if (r0 + M - 1 < ns) X[r0 + M - 1] = sp;
}

// first grid index i with k0 p_i >= v (N if none): the closed form
// i = N log(v / k0) / G, G = log(kmax / kmin), settled against the table
__device__ __forceinline__ int grid_lower_index(const ce_grid_dev &g, double v)
{
    gdouble *p = as_global(g.p);
    const double k0 = g.k0;
    const int N = g.N;
    const double t = (double)N * log(v / k0) / g.G;
    int i = (t > 0.0) ? (t < (double)N ? (int)ceil(t) : N) : 0;
    // both neighbours are loaded at once; a walk follows only for a poor guess
    const double below = k0 * p[max(i - 1, 0)], at = k0 * p[min(i, N - 1)];
    if ((i > 0 && below >= v) || (i < N && at < v)) {
        while (i > 0 && k0 * p[i - 1] >= v) --i;
        while (i < N && k0 * p[i] < v) ++i;
    }
    return i;
}

// interval j of the LDS-staged nodes xs[0..cnt) with xs[j] <= v < xs[j+1] (the
// largest j <= cnt - 2 with xs[j] <= v; 0 below xs[0]), for CE_PTS_PER values at
// once: binary lifting with a fixed step sequence, so the loads of all values
// go out together at every step (cnt <= CE_CHUNK_NODES <= 1025)
#define CE_PTS_PER 8
static_assert(CE_CHUNK_NODES <= 1025, "interval search steps");
__device__ __forceinline__ void lds_intervals(const double *xs, int cnt, const double (&v)[CE_PTS_PER],
                                              int (&j)[CE_PTS_PER])
{
#pragma unroll
    for (int t = 0; t < CE_PTS_PER; ++t) j[t] = 0;
#pragma unroll
    for (int step = 512; step > 0; step >>= 1) {
#pragma unroll
        for (int t = 0; t < CE_PTS_PER; ++t) {
            const int c = min(j[t] + step, cnt - 2);
            if (xs[c] <= v[t]) j[t] = c;
        }
    }
}

struct chunk_geom {
    int a, b;        // owned intervals [a, b) (b = n - 1 on the last chunk)
    int lo, hi;      // staged nodes [lo, hi]
    int u0, u1;      // unknown nodes of the truncated system
};

__device__ __forceinline__ chunk_geom make_geom(const ce_chunk_dev &ch, int n)
{
    chunk_geom g;
    g.a = ch.a;
    g.b = ch.b;
    g.u0 = max(1, g.a - CE_SPLINE_HALO);
    g.u1 = min(n - 2, g.b + CE_SPLINE_HALO);
    g.lo = g.u0 - 1;
    g.hi = g.u1 + 1;
    return g;
}

// right-hand side of the interpolating spline, 3 (slope_{r+1} - slope_r), from the
// lane's own values yl[q] = y at staged index lane M + q (q <= M + 1, in registers);
// the slope of the row's left interval is carried over from the previous row
struct spline_rhs {
    const double *yl;
    double sprev;
    __device__ __forceinline__ double operator()(int j, int, double hm, double hp)
    {
        const double sm = (j == 0) ? (hm != 0.0 ? (yl[1] - yl[0]) * rcp_nr(hm) : 0.0) : sprev;
        const double sp = hp != 0.0 ? (yl[j + 2] - yl[j + 1]) * rcp_nr(hp) : 0.0;
        sprev = sp;
        return 3.0 * (sp - sm);
    }
};

// k_resample: S(x_i) for the resample points whose interval lies in [a, b); one
// wavefront per chunk
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4)))
k_resample(const ce_chunk_dev *__restrict__ chunks,
                                                 const ce_spec_dev *__restrict__ specs,
                                                 const ce_grid_dev *__restrict__ grids,
                                                 const double *__restrict__ k_all,
                                                 const double *__restrict__ pk_all, double *__restrict__ ws_all,
                                                 int32_t *__restrict__ ranges)
{
    // two LDS arrays (nodes and spline coefficients; the values stay in global memory
    // and registers), 10.4 KiB: 15 chunks fit a CU
    __shared__ double lds[2 * CE_CHUNK_NODES];
    const ce_chunk_dev ch = chunks[blockIdx.x];
    const ce_grid_dev &g = grids[ch.grid];
    gdouble *gp = as_global(g.p);
    const double gk0 = g.k0;
    const int n = ch.n_in;
    const double *k = k_all + ch.in_off;
    double *yres = ws_all + ch.ws_off;
    const chunk_geom G = make_geom(ch, n);
    gdouble *ys = as_global(pk_all + ch.in_off + G.lo);   // y at staged index i
    double *xs = lds, *D = xs + CE_CHUNK_NODES;
    const int lane = threadIdx.x;
    const int cnt = G.hi - G.lo + 1;
    // the staged nodes go to registers first, so that their loads overlap the search
    // for the resample range and the loads of the first resample points; each lane
    // also loads the values of its own rows of the spline system
    constexpr int SPL = (CE_CHUNK_NODES + 63) / 64;
    double kv[SPL];
#pragma unroll
    for (int u = 0; u < SPL; ++u) {
        const int i = lane + 64 * u;
        if (i < cnt) kv[u] = k[G.lo + i];
    }
    double yl[CE_SOLVE_M + 2];
#pragma unroll
    for (int q = 0; q < CE_SOLVE_M + 2; ++q) yl[q] = ys[min(lane * CE_SOLVE_M + q, cnt - 1)];
    // resample index ranges: lanes 0, 1 bound the points of the owned intervals
    // [a, b), lanes 2, 3 those of the staged intervals [lo, hi) (kept for the adjoint)
    int ib;
    {
        const int q = lane & 3;
        const int node = q == 0 ? G.a : (q == 1 ? G.b : (q == 2 ? G.lo : G.hi));
        const bool edge = (q & 1) ? (node == n - 1) : (node == 0);
        ib = edge ? ((q & 1) ? g.N : 0) : grid_lower_index(g, k[node]);
    }
    if (ranges && lane < 4) ranges[4 * ch.id + lane] = ib;
    const int i0 = __shfl(ib, 0), i1 = __shfl(ib, 1);
    const double klo = k[0], khi = k[n - 1];
    // loads at clamped indices, unconditionally, so that they all go out at once
    const int ilast = max(i1 - 1, 0);
    double pq[CE_PTS_PER];   // the raw table values: nothing waits on them before use
#pragma unroll
    for (int t = 0; t < CE_PTS_PER; ++t) pq[t] = gp[min(i0 + lane + 64 * t, ilast)];
#pragma unroll
    for (int u = 0; u < SPL; ++u) {
        const int i = lane + 64 * u;
        if (i < cnt) xs[i] = kv[u];
    }
    wave_sync();
    const int ns = G.u1 - G.u0 + 1;
    wave_spline_solve(xs, ns, spline_rhs{yl, 0.0}, D);
    wave_sync();
    // c at staged node li: D[li - 1] for u0 <= node <= u1, 0 at the global ends
    auto cval = [&](int li) {
        const int node = G.lo + li;
        return (node < 1 || node > n - 2) ? 0.0 : D[node - G.u0];
    };
    for (int base = i0 + lane; base < i1; base += 64 * CE_PTS_PER) {
        int jv[CE_PTS_PER];
        if (base != i0 + lane) {
#pragma unroll
            for (int t = 0; t < CE_PTS_PER; ++t) pq[t] = gp[min(base + 64 * t, ilast)];
        }
        double xq[CE_PTS_PER];
#pragma unroll
        for (int t = 0; t < CE_PTS_PER; ++t) xq[t] = gk0 * pq[t];
        lds_intervals(xs, cnt, xq, jv);
        double y0[CE_PTS_PER], y1[CE_PTS_PER];
#pragma unroll
        for (int t = 0; t < CE_PTS_PER; ++t) {
            y0[t] = ys[jv[t]];
            y1[t] = ys[jv[t] + 1];
        }
#pragma unroll
        for (int t = 0; t < CE_PTS_PER; ++t) {
            const int i = base + 64 * t;
            if (i >= i1) break;
            const int j = jv[t];
            double S;
            if (!(xq[t] >= klo && xq[t] <= khi)) {
                S = __builtin_nan("");
            } else {
                const double dx = xs[j + 1] - xs[j], rdx = rcp_nr(dx);
                const double dy = y1[t] - y0[t];
                const double delx = xq[t] - xs[j];
                const double cj = cval(j), cj1 = cval(j + 1);
                const double bb = dy * rdx - dx * (cj1 + 2.0 * cj) / 3.0;
                const double dd = (cj1 - cj) * rdx / 3.0;
                S = y0[t] + delx * (bb + delx * (cj + delx * dd));
            }
            yres[i] = S;
        }
    }
}

// k_prepare: y_i = bias_q(i) window(x_i) S(x_i) -> r2c -> conj, per spectrum and q'
__global__ void __launch_bounds__(1024) k_prepare(const ce_spec_dev *__restrict__ specs,
                                                  const ce_grid_dev *__restrict__ grids, int nq,
                                                  const double *__restrict__ ws_all, cd *__restrict__ xf_all,
                                                  int max_M)
{
    extern __shared__ cd smem[];
    const ce_spec_dev sp = specs[blockIdx.x];
    const ce_grid_dev &g = grids[sp.grid];
    const int M = g.M;
    const uint32_t plan = g.rp.code;
    const int T = blockDim.x, tid = threadIdx.x;
    // S(x_i) as pairs (S_2j, S_2j+1): complex slot j of the transform input
    gcd *ypair = as_global(reinterpret_cast<const cd *>(ws_all + sp.ws_off));
    gcd *tw = as_global(reinterpret_cast<const cd *>(g.tw));
    cd *lt = smem + lds_data_len(max_M);
    load_twiddles(lt, M, tw);
    for (int q = 0; q < nq; ++q) {
        gcd *bwpair = as_global(reinterpret_cast<const cd *>(g.bw[q]));
        // four rounds of loads in flight at a time
        for (int j0 = 0; j0 * T < M; j0 += 4) {
            cd bv[4], yv[4];
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const int j = min(tid + (j0 + h) * T, M - 1);
                bv[h] = ldg(bwpair + j);
                yv[h] = ldg(ypair + j);
            }
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const int j = tid + (j0 + h) * T;
                if (j < M) smem[pidx(j)] = mk(bv[h].x * yv[h].x, bv[h].y * yv[h].y);
            }
        }
        lds_barrier();
        fft_lds(smem, M, plan, lt);
        gcd_out *xf = as_global_out(xf_all + sp.xf_off + (int64_t)q * (M + 1));
        for (int it0 = 0; it0 * T <= M / 2; it0 += 3) {
            cd w1[3], w2[3];
#pragma unroll
            for (int h = 0; h < 3; ++h) {
                const int kk = min(tid + (it0 + h) * T, M / 2);
                w1[h] = ldg(tw + kk);
                w2[h] = mk(-w1[h].x, w1[h].y);   // tw[M - kk] = -conj(tw[kk])
            }
#pragma unroll
            for (int h = 0; h < 3; ++h) {
                const int k1 = tid + (it0 + h) * T, k2 = M - k1;
                if (k1 <= M / 2) {
                    const cd Z1 = smem[pidx(k1)];
                    const cd Z2 = smem[pidx(k2 == M ? 0 : k2)];
                    stg(xf + k1, cconj(r2c_split(Z1, Z2, w1[h])));
                    if (k2 != k1) stg(xf + k2, cconj(r2c_split(Z2, Z1, w2[h])));
                }
            }
        }
        lds_barrier();
    }
}

// Akima (1970, non-rounded ends) value on interval [X2, X3] of the 6-point stencil
// X[0..5], Y[0..5] (GSL akima: slopes m, weights alpha).  With dY non-null the
// derivative of the value with respect to Y[0..5] is returned too.
template <bool WANT_DY>
__device__ __forceinline__ double akima_stencil(const double *X, const double *Y, double xq, double *dY)
{
    double m[5], inv[5];
#pragma unroll
    for (int a = 0; a < 5; ++a) {
        inv[a] = 1.0 / (X[a + 1] - X[a]);
        m[a] = (Y[a + 1] - Y[a]) * inv[a];
    }
    const double h = X[3] - X[2];
    const double delx = xq - X[2];
    const double A = fabs(m[1] - m[0]), B = fabs(m[3] - m[2]);
    const double ne = B + A;
    double val;
    double dm[5] = {0, 0, 0, 0, 0};
    double dY2 = 1.0;
    if (ne == 0.0) {
        val = Y[2] + delx * m[2];
        dm[2] = delx;
    } else {
        const double C = fabs(m[2] - m[1]), D = fabs(m[4] - m[3]);
        const double nen = D + C;
        const double alpha = A / ne;
        double tn, beta = 0.0;
        if (nen == 0.0) {
            tn = m[2];
        } else {
            beta = C / nen;
            tn = (1.0 - beta) * m[2] + beta * m[3];
        }
        const double bb = (1.0 - alpha) * m[1] + alpha * m[2];
        const double cc = (3.0 * m[2] - 2.0 * bb - tn) / h;
        const double dd = (bb + tn - 2.0 * m[2]) / (h * h);
        val = Y[2] + delx * (bb + delx * (cc + dd * delx));
        if (WANT_DY) {
            const double u = delx / h;
            const double dv_db = h * (u - 2 * u * u + u * u * u);
            const double dv_dt = h * (u * u * u - u * u);
            const double dv_dm2 = h * (3 * u * u - 2 * u * u * u);
            dm[2] += dv_dm2;
            // b = m1 + alpha (m2 - m1)
            dm[1] += dv_db * (1.0 - alpha);
            dm[2] += dv_db * alpha;
            const double dv_dalpha = dv_db * (m[2] - m[1]);
            const double dalpha_dA = B / (ne * ne), dalpha_dB = -A / (ne * ne);
            const double sA = (m[1] > m[0]) - (m[1] < m[0]);
            const double sB = (m[3] > m[2]) - (m[3] < m[2]);
            dm[1] += dv_dalpha * dalpha_dA * sA;
            dm[0] -= dv_dalpha * dalpha_dA * sA;
            dm[3] += dv_dalpha * dalpha_dB * sB;
            dm[2] -= dv_dalpha * dalpha_dB * sB;
            if (nen == 0.0) {
                dm[2] += dv_dt;
            } else {
                dm[2] += dv_dt * (1.0 - beta);
                dm[3] += dv_dt * beta;
                const double dv_dbeta = dv_dt * (m[3] - m[2]);
                const double dbeta_dC = D / (nen * nen), dbeta_dD = -C / (nen * nen);
                const double sC = (m[2] > m[1]) - (m[2] < m[1]);
                const double sD = (m[4] > m[3]) - (m[4] < m[3]);
                dm[2] += dv_dbeta * dbeta_dC * sC;
                dm[1] -= dv_dbeta * dbeta_dC * sC;
                dm[4] += dv_dbeta * dbeta_dD * sD;
                dm[3] -= dv_dbeta * dbeta_dD * sD;
            }
        }
    }
    if (WANT_DY) {
#pragma unroll
        for (int a = 0; a < 6; ++a) dY[a] = 0.0;
        dY[2] = dY2;
#pragma unroll
        for (int a = 0; a < 5; ++a) {
            dY[a + 1] += dm[a] * inv[a];
            dY[a] -= dm[a] * inv[a];
        }
    }
    return val;
}

// table value i of the c2r output held (as conj of the forward transform) in LDS
__device__ __forceinline__ double c2r_value(const cd *s, int i)
{
    const cd z = s[pidx(i >> 1)];
    return (i & 1) ? -z.y : z.x;
}

struct table_args {
    const ce_unit_dev *units;
    const ce_spec_dev *specs;
    const ce_grid_dev *grids;
    const ce_table_desc *tabs;
    int ntab;
    const cd *xf;
    int nsep;
    const int32_t *sidx;   // [grid][nsep] FFT interval index
    const double *sep;
    double *itab;          // [spec][tab][nsep]
    double *full;          // [unit][full_stride] or null
    int64_t full_stride;
    const double *A;       // [nz][nl][ntab][nsep] (adjoint only)
    const double *gout;    // [spec][nz][nl][nsep] (adjoint only)
    int nz, nl;
    cd *xbar;              // adjoint output, per spec ntab*(M+1) at specs[].xb_off
    double *wbuf;          // [unit][nsep][6] scratch
    int max_M;
    const int4 *packs;     // per workgroup: first unit, unit count, groups (null: one unit each)
    int lt_off;            // LDS offset (complex) of the twiddle tables, after the groups' data
};

__global__ void __launch_bounds__(1024) k_tables(table_args a)
{
    extern __shared__ cd smem[];
    // the workgroup's units: `count` units of one transform length, one per group of
    // blockDim.x / P threads with its own LDS region (groups past count repeat the
    // first unit and store nothing, so that every group passes the same barriers)
    int first = blockIdx.x, count = 1, P = 1;
    if (a.packs) {
        const int4 pk = a.packs[blockIdx.x];
        first = pk.x;
        count = pk.y;
        P = pk.z;
    }
    const int T = blockDim.x / P;
    // groups are whole wavefronts: the group index (and all that follows from it) is
    // uniform across the wavefront
    const int grp = __builtin_amdgcn_readfirstlane(threadIdx.x / T);
    const int tid = threadIdx.x - grp * T;
    const bool live = grp < count;
    const int ui = first + (live ? grp : 0);
    const ce_unit_dev u = a.units[ui];
    const ce_spec_dev sp = a.specs[u.spec];
    const ce_grid_dev &g = a.grids[sp.grid];
    const ce_table_desc td = a.tabs[u.tab];
    const int M = g.M, N = g.N;
    const uint32_t plan = g.rp.code;
    const double r0 = g.r0;
    gcd *tw = as_global(reinterpret_cast<const cd *>(g.tw));
    gcd *U = as_global(reinterpret_cast<const cd *>(g.U[u.tab]));
    gdouble *os = as_global(g.os[u.tab]);
    gdouble *gp = as_global(g.p);
    gcd *xf = as_global(a.xf + sp.xf_off + (int64_t)td.qidx * (M + 1));
    gdouble *sep = as_global(a.sep);
    const int32_t *sidx = a.sidx + (int64_t)sp.grid * a.nsep;
    const bool adj = a.gout != nullptr;
    cd *lt = smem + a.lt_off;
    load_twiddles(lt, M, tw);   // by the whole workgroup: its groups share M
    cd *sm = smem + grp * lds_data_len(M);
    // the interval of this thread's first separation, fetched ahead of the transform
    const bool inreg = a.nsep <= 2 * T;
    const int r_a = tid;
    const bool has_a = inreg && r_a < a.nsep;
    const int j_a = has_a ? sidx[r_a] : 2;
    // the cotangent of I_t(r) in adjoint mode, sum over (z, l) of A gout (in that order),
    // from (z, l) pair q0 on, four pairs' loads in flight at a time
    gdouble *gA = as_global(a.A), *gG = as_global(a.gout);
    const int nzl = a.nz * a.nl;
    auto zl_index = [&](int q, int r, int64_t &ia, int64_t &ig) {
        const int z = q / a.nl, l = q - (q / a.nl) * a.nl;
        ia = (((int64_t)z * a.nl + l) * a.ntab + u.tab) * a.nsep + r;
        ig = (((int64_t)u.spec * a.nz + z) * a.nl + l) * a.nsep + r;
    };
    auto cotangent_sum = [&](int r, int q0, double ib) {
        for (; q0 < nzl; q0 += 4) {
            double av[4], gv[4];
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                int64_t ia, ig;
                zl_index(min(q0 + h, nzl - 1), r, ia, ig);
                av[h] = gA[ia];
                gv[h] = gG[ig];
            }
#pragma unroll
            for (int h = 0; h < 4; ++h)
                if (q0 + h < nzl) ib += av[h] * gv[h];
        }
        return ib;
    };
    // kernel multiply + c2r merge (stored conjugated: the forward FFT then gives conj(z)),
    // over the pairs (k, M - k), three rounds of loads in flight at a time; the merge
    // twiddle of M - k is -conj of that of k (N = 2M), exactly so for the correctly
    // rounded table
    for (int it0 = 0; it0 * T <= M / 2; it0 += 3) {
        cd x1[3], x2[3], u1[3], u2[3], w1[3];
#pragma unroll
        for (int h = 0; h < 3; ++h) {
            const int kk = min(tid + (it0 + h) * T, M / 2);
            x1[h] = ldg(xf + kk);
            u1[h] = ldg(U + kk);
            x2[h] = ldg(xf + (M - kk));
            u2[h] = ldg(U + (M - kk));
            w1[h] = ldg(tw + kk);
        }
#pragma unroll
        for (int h = 0; h < 3; ++h) {
            const int k1 = tid + (it0 + h) * T, k2 = M - k1;
            if (k1 <= M / 2) {
                cd X1 = cmul(x1[h], u1[h]);
                cd X2 = cmul(x2[h], u2[h]);
                if (k1 == 0) { X1.y = 0.0; X2.y = 0.0; }
                sm[pidx(k1)] = cconj(c2r_merge(X1, X2, w1[h]));
                if (k2 != k1 && k2 < M) sm[pidx(k2)] = cconj(c2r_merge(X2, X1, mk(-w1[h].x, w1[h].y)));
            }
        }
    }
    lds_barrier();
    fft_lds(sm, M, plan, lt, tid, T);
    if (a.full) {
        gdouble_out *out = as_global_out(a.full + (int64_t)ui * a.full_stride);
        if (live)
            for (int i = tid; i < N; i += T) out[i] = os[i] * c2r_value(sm, i);
    }
    // Akima at separation r from gathered operands; in adjoint mode also the
    // cotangents of the six stencil values of the c2r output (value derivative *
    // cotangent of I_t(r) * output scale)
    auto akima_finish = [&](int r, int j, const double *X, const double *O, double ib, double xq, double *w) {
        double Y[6], dY[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) Y[m] = O[m] * c2r_value(sm, j - 2 + m);
        const double v = adj ? akima_stencil<true>(X, Y, xq, dY) : akima_stencil<false>(X, Y, xq, dY);
        if (a.itab && live) a.itab[((int64_t)u.spec * a.ntab + u.tab) * a.nsep + r] = v;
        if (adj) {
#pragma unroll
            for (int m = 0; m < 6; ++m) w[m] = ib * dY[m] * O[m];
        }
    };
    auto akima_at = [&](int r, double *w) -> int {
        const int j = sidx[r];
        double X[6], O[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            X[m] = r0 * gp[j - 2 + m];
            O[m] = os[j - 2 + m];
        }
        double ib = 0.0;
        if (adj) ib = cotangent_sum(r, 0, ib);
        akima_finish(r, j, X, O, ib, sep[r], w);
        return j;
    };
    // up to two separations per thread keep their cotangents in registers across the
    // LDS clear; more separations go through the global scratch wbuf
    double w0[6], w1[6];
    int j0 = -1, j1 = -1;
    double *wb = a.wbuf + (int64_t)ui * a.nsep * 6;
    if (inreg) {
        if (has_a) {
            double X_a[6], O_a[6];
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                X_a[m] = r0 * gp[j_a - 2 + m];
                O_a[m] = os[j_a - 2 + m];
            }
            const double ib_a = adj ? cotangent_sum(r_a, 0, 0.0) : 0.0;
            akima_finish(r_a, j_a, X_a, O_a, ib_a, sep[r_a], w0);
            j0 = j_a;
        }
        if (tid + T < a.nsep) j1 = akima_at(tid + T, w1);
    } else {
#pragma unroll 2
        for (int r = tid; r < a.nsep; r += T) {
            double w[6];
            akima_at(r, w);
            if (adj && live) {
#pragma unroll
                for (int m = 0; m < 6; ++m) wb[r * 6 + m] = w[m];
            }
        }
    }
    if (!adj) return;
    lds_barrier();
    for (int i = tid; i < N; i += T) lds_real(sm, i) = 0.0;
    lds_barrier();
    if (inreg) {
        if (j0 >= 0) {
#pragma unroll
            for (int m = 0; m < 6; ++m) atomicAdd(&lds_real(sm, j0 - 2 + m), w0[m]);
        }
        if (j1 >= 0) {
#pragma unroll
            for (int m = 0; m < 6; ++m) atomicAdd(&lds_real(sm, j1 - 2 + m), w1[m]);
        }
    } else {
        for (int r = tid; r < a.nsep; r += T) {
            const int j = sidx[r];
#pragma unroll
            for (int m = 0; m < 6; ++m) atomicAdd(&lds_real(sm, j - 2 + m), live ? wb[r * 6 + m] : 0.0);
        }
    }
    lds_barrier();
    // transposed c2r: Xbar_k = r2c(ybar)_k, doubled for 0 < k < M, real at k = 0, M
    fft_lds(sm, M, plan, lt, tid, T);
    gcd_out *xb = as_global_out(a.xbar + sp.xb_off + (int64_t)u.tab * (M + 1));
    for (int it0 = 0; it0 * T <= M / 2; it0 += 3) {
        cd u1[3], u2[3], w1[3], w2[3];
#pragma unroll
        for (int h = 0; h < 3; ++h) {
            const int kk = min(tid + (it0 + h) * T, M / 2);
            u1[h] = ldg(U + kk);
            u2[h] = ldg(U + (M - kk));
            w1[h] = ldg(tw + kk);
            w2[h] = mk(-w1[h].x, w1[h].y);   // tw[M - kk] = -conj(tw[kk])
        }
#pragma unroll
        for (int h = 0; h < 3; ++h) {
            const int k1 = tid + (it0 + h) * T, k2 = M - k1;
            if (k1 <= M / 2 && live) {
                const cd Z1 = sm[pidx(k1)];
                const cd Z2 = sm[pidx(k2 == M ? 0 : k2)];
                cd R1 = r2c_split(Z1, Z2, w1[h]);
                cd R2 = r2c_split(Z2, Z1, w2[h]);
                R1 = (k1 == 0) ? mk(R1.x, 0.0) : cscale(R1, 2.0);
                R2 = (k2 == M) ? mk(R2.x, 0.0) : cscale(R2, 2.0);
                // X = xf * U  ->  xf_bar = conj(U) * X_bar
                stg(xb + k1, cmul(cconj(u1[h]), R1));
                if (k2 != k1) stg(xb + k2, cmul(cconj(u2[h]), R2));
            }
        }
    }
}

__global__ void k_project(int nspec, int ntab, int nz, int nl, int nsep, const double *__restrict__ A,
                          const double *__restrict__ itab, double *__restrict__ out)
{
    const int64_t total = (int64_t)nspec * nz * nl * nsep;
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(idx % nsep);
        const int64_t rest = idx / nsep;
        const int l = (int)(rest % nl);
        const int z = (int)((rest / nl) % nz);
        const int64_t s = rest / ((int64_t)nl * nz);
        double acc = 0.0;
        for (int t = 0; t < ntab; ++t)
            acc += A[(((int64_t)z * nl + l) * ntab + t) * nsep + r] * itab[((int64_t)s * ntab + t) * nsep + r];
        out[idx] = acc;
    }
}

// k_adj_c2r: per spectrum, Ybar = conj(sum_t xbar_t) -> transposed r2c (a c2r with
// the interior halved) -> times bias*window -> sbar (cotangent of S(x_i))
// the multipoles of one spectrum, out[z][l][r] = sum_t A[z][l][t][r] itab[t][r] (the
// projection of k_project), folded into k_adj_c2r in the composite adjoint call
struct proj_args {
    const double *A;       // [nz][nl][ntab][nsep]
    const double *itab;    // [spec][ntab][nsep]
    double *out;           // [spec][nz][nl][nsep], or null: no projection
    int nz, nl, nsep;
};

__global__ void __launch_bounds__(1024) k_adj_c2r(const ce_spec_dev *__restrict__ specs,
                                                  const ce_grid_dev *__restrict__ grids,
                                                  const ce_table_desc *__restrict__ tabs, int ntab, int nq,
                                                  double *__restrict__ ws_all, const cd *__restrict__ xbar_all,
                                                  int max_M, proj_args pj)
{
    extern __shared__ cd smem[];
    if (pj.out) {
        const int64_t s = blockIdx.x;
        const int nout = pj.nz * pj.nl * pj.nsep;
        gdouble *A = as_global(pj.A);
        gdouble *it = as_global(pj.itab + s * ntab * pj.nsep);
        gdouble_out *o = as_global_out(pj.out + s * nout);
        for (int e = threadIdx.x; e < nout; e += blockDim.x) {
            const int r = e % pj.nsep, zl = e / pj.nsep;
            double acc = 0.0;
            for (int t = 0; t < ntab; ++t) acc += A[((int64_t)zl * ntab + t) * pj.nsep + r] * it[t * pj.nsep + r];
            o[e] = acc;
        }
    }
    const ce_spec_dev sp = specs[blockIdx.x];
    const ce_grid_dev &g = grids[sp.grid];
    const int M = g.M, N = g.N;
    const uint32_t plan = g.rp.code;
    const int T = blockDim.x, tid = threadIdx.x;
    // the cotangent of S(x_i) as pairs (2j, 2j+1)
    gcd_out *sbar = as_global_out(reinterpret_cast<cd *>(ws_all + sp.ws_off + N));
    gcd *tw = as_global(reinterpret_cast<const cd *>(g.tw));
    cd *lt = smem + lds_data_len(max_M);
    load_twiddles(lt, M, tw);
    gcd *xbar = as_global(xbar_all + sp.xb_off);
    for (int q = 0; q < nq; ++q) {
        // the tables' cotangent spectra summed per pair (k, M - k), two rounds of loads
        // in flight at a time
        for (int it0 = 0; it0 * T <= M / 2; it0 += 2) {
            cd S1[2], S2[2], w1[2], w2[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int kk = min(tid + (it0 + h) * T, M / 2);
                w1[h] = ldg(tw + kk);
                w2[h] = mk(-w1[h].x, w1[h].y);   // tw[M - kk] = -conj(tw[kk])
                S1[h] = mk(0.0, 0.0);
                S2[h] = mk(0.0, 0.0);
            }
            for (int t = 0; t < ntab; ++t) {
                if (tabs[t].qidx != q) continue;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int kk = min(tid + (it0 + h) * T, M / 2);
                    S1[h] = cadd(S1[h], ldg(xbar + ((int64_t)t * (M + 1) + kk)));
                    S2[h] = cadd(S2[h], ldg(xbar + ((int64_t)t * (M + 1) + M - kk)));
                }
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int k1 = tid + (it0 + h) * T, k2 = M - k1;
                if (k1 <= M / 2) {
                    cd Y1 = cconj(S1[h]), Y2 = cconj(S2[h]);
                    Y1 = (k1 == 0) ? mk(Y1.x, 0.0) : cscale(Y1, 0.5);
                    Y2 = (k2 == M) ? mk(Y2.x, 0.0) : cscale(Y2, 0.5);
                    smem[pidx(k1)] = cconj(c2r_merge(Y1, Y2, w1[h]));
                    if (k2 != k1 && k2 < M) smem[pidx(k2)] = cconj(c2r_merge(Y2, Y1, w2[h]));
                }
            }
        }
        lds_barrier();
        fft_lds(smem, M, plan, lt);
        gcd *bwpair = as_global(reinterpret_cast<const cd *>(g.bw[q]));
        for (int j0 = 0; j0 * T < M; j0 += 4) {
            cd bv[4], pv[4];
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const int j = min(tid + (j0 + h) * T, M - 1);
                bv[h] = ldg(bwpair + j);
                pv[h] = (q == 0) ? mk(0.0, 0.0) : mk(sbar[j].x, sbar[j].y);
            }
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const int j = tid + (j0 + h) * T;
                if (j < M) {
                    // c2r outputs 2j, 2j+1 of slot j (the transform holds their conjugate)
                    const cd z = smem[pidx(j)];
                    stg(sbar + j, mk(bv[h].x * z.x + pv[h].x, bv[h].y * (-z.y) + pv[h].y));
                }
            }
        }
        lds_barrier();
    }
}

// k_spline_adj: transposed resample and spline on the chunk's nodes -> gpk[a..b)
// (the last chunk also owns node n - 1).  The direct and c-coefficient cotangents
// of every resample point are scattered with LDS atomics, lambda = T^{-1} cbar is
// solved on the truncated system, and dP_j collects the y-dependence of the
// right-hand side g (twofast.c:399-401 / the natural-spline equations).
__global__ void __launch_bounds__(64) k_spline_adj(const ce_chunk_dev *__restrict__ chunks,
                                                   const ce_spec_dev *__restrict__ specs,
                                                   const ce_grid_dev *__restrict__ grids,
                                                   const double *__restrict__ k_all,
                                                   const double *__restrict__ ws_all, double *__restrict__ gpk_all,
                                                   const int32_t *__restrict__ ranges)
{
    __shared__ double lds[3 * CE_CHUNK_NODES];
    const ce_chunk_dev ch = chunks[blockIdx.x];
    const ce_grid_dev &g = grids[ch.grid];
    gdouble *gp = as_global(g.p);
    const double gk0 = g.k0;
    const int n = ch.n_in, N = g.N;
    const double *k = k_all + ch.in_off;
    const double *sbar = ws_all + ch.ws_off + N;
    double *gpk = gpk_all + ch.in_off;
    const chunk_geom G = make_geom(ch, n);
    const int cnt = G.hi - G.lo + 1;
    const int ns = G.u1 - G.u0 + 1;
    double *xs = lds, *pb = xs + CE_CHUNK_NODES, *cb = pb + CE_CHUNK_NODES;
    const int lane = threadIdx.x;
    // staged nodes to registers first (their loads overlap the range search)
    constexpr int SPL = (CE_CHUNK_NODES + 63) / 64;
    double kv[SPL];
#pragma unroll
    for (int u = 0; u < SPL; ++u) {
        const int i = lane + 64 * u;
        if (i < cnt) kv[u] = k[G.lo + i];
    }
    // resample points with intervals in the staged range [lo, hi), as k_resample of
    // the same prepare stage found them
    // (clamped to the grid, so that a stage run out of order stays in bounds)
    const int i1 = min(max(ranges[4 * ch.id + 3], 0), N);
    const int i0 = min(max(ranges[4 * ch.id + 2], 0), i1);
    const double klo = k[0], khi = k[n - 1];
    // loads at clamped indices, unconditionally, so that they all go out at once
    const int ilast = max(i1 - 1, 0);
    double xq[CE_PTS_PER], sv[CE_PTS_PER];
#pragma unroll
    for (int t = 0; t < CE_PTS_PER; ++t) {
        const int i = min(i0 + lane + 64 * t, ilast);
        xq[t] = gk0 * gp[i];
        sv[t] = sbar[i];
    }
#pragma unroll
    for (int u = 0; u < SPL; ++u) {
        const int i = lane + 64 * u;
        if (i < cnt) {
            xs[i] = kv[u];
            pb[i] = 0.0;
            cb[i] = 0.0;
        }
    }
    wave_sync();
    // scatter the resample cotangents of the points with intervals in [lo, hi)
    for (int base = i0 + lane; base < i1; base += 64 * CE_PTS_PER) {
        int jv[CE_PTS_PER];
        if (base != i0 + lane) {
#pragma unroll
            for (int t = 0; t < CE_PTS_PER; ++t) {
                const int i = min(base + 64 * t, ilast);
                xq[t] = gk0 * gp[i];
                sv[t] = sbar[i];
            }
        }
        lds_intervals(xs, cnt, xq, jv);
#pragma unroll
        for (int t = 0; t < CE_PTS_PER; ++t) {
            if (base + 64 * t >= i1) break;
            if (!(xq[t] >= klo && xq[t] <= khi)) continue;
            const int j = jv[t];
            const double dx = xs[j + 1] - xs[j], rdx = rcp_nr(dx);
            const double d = xq[t] - xs[j];
            const double d3 = d * d * d * rdx / 3.0;
            atomicAdd(&pb[j], sv[t] * (1.0 - d * rdx));
            atomicAdd(&pb[j + 1], sv[t] * d * rdx);
            atomicAdd(&cb[j], sv[t] * (-2.0 * d * dx / 3.0 + d * d - d3));
            atomicAdd(&cb[j + 1], sv[t] * (-d * dx / 3.0 + d3));
        }
    }
    wave_sync();
    // the transposed solve: the spline matrix is symmetric; node u0 + r sits at staged
    // index r + 1.  The solution overwrites cb in place (D[r] = cb[r]): the one
    // wavefront has read every right-hand side before its first store
    double *D = cb;
    wave_spline_solve(xs, ns, [&](int, int r, double, double) { return cb[r + 1]; }, D);
    wave_sync();
    auto lam = [&](int node) { return (node < G.u0 || node > G.u1) ? 0.0 : D[node - G.u0]; };
    const int jend = (G.b == n - 1) ? n : G.b;
    for (int jj = G.a + lane; jj < jend; jj += 64) {
        const int li = jj - G.lo;
        double acc = pb[li];
        const int sys = n - 2;
        // g_s (node s + 1) = 3((y_{s+2} - y_{s+1})/h_{s+1} - (y_{s+1} - y_s)/h_s)
        const double xl = xs[max(li - 1, 0)], xc = xs[li], xr = xs[min(li + 1, cnt - 1)];
        const double gm = (jj >= 1) ? 3.0 * rcp_nr(xc - xl) : 0.0;
        const double gp = (jj <= n - 2) ? 3.0 * rcp_nr(xr - xc) : 0.0;
        if (jj - 2 >= 0 && jj - 2 < sys) acc += lam(jj - 1) * gm;
        if (jj - 1 >= 0 && jj - 1 < sys) acc += lam(jj) * (-gp - gm);
        if (jj < sys) acc += lam(jj + 1) * gp;
        gpk[jj] = acc;
    }
}

}  // namespace

// ---------------------------------------------------------------------------------
extern "C" {

// raise a kernel's dynamic-LDS limit once per (kernel, size): the attribute call is a
// host round trip, kept out of the per-step launch path
static int set_lds(const void *fn, size_t bytes)
{
    static std::mutex mu;
    static const void *fns[8] = {nullptr};
    static size_t done[8] = {0};
    std::lock_guard<std::mutex> g(mu);
    int slot = -1;
    for (int i = 0; i < 8; ++i) {
        if (fns[i] == fn) { slot = i; break; }
        if (!fns[i]) { fns[i] = fn; slot = i; break; }
    }
    if (slot >= 0 && done[slot] >= bytes) return 0;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess && slot >= 0) done[slot] = bytes;
    return e == hipSuccess ? 0 : (int)e;
}

int ce_launch_grid_tables(void *stream, int N, double kmin, double kmax, double k0, double r0, const double *d_p,
                          int ntab, const ce_table_desc *tabs, const double *tab_qnu, int nq, const double *qprime,
                          double *const *d_U, double *const *d_os, double *const *d_bw)
{
    grid_tab_args a;
    a.N = N;
    a.kmin = kmin;
    a.kmax = kmax;
    a.k0 = k0;
    a.r0 = r0;
    a.p = d_p;
    a.ntab = ntab;
    a.nq = nq;
    for (int t = 0; t < ntab; ++t) {
        a.n[t] = tabs[t].n;
        a.l[t] = tabs[t].l;
        a.qidx[t] = tabs[t].qidx;
        a.rpow[t] = tabs[t].rpow;
        a.qnu[t] = tab_qnu[t];
        a.U[t] = d_U[t];
        a.os[t] = d_os[t];
    }
    for (int q = 0; q < nq; ++q) {
        a.qprime[q] = qprime[q];
        a.bw[q] = d_bw[q];
    }
    const int threads = 256;
    const int blocks = (N + threads - 1) / threads;
    hipLaunchKernelGGL(k_grid_tables, dim3(blocks), dim3(threads), 0, (hipStream_t)stream, a);
    return (int)hipGetLastError();
}

int ce_launch_resample(void *stream, int nchunk, const ce_chunk_dev *d_chunks, const ce_spec_dev *d_spec,
                       const ce_grid_dev *d_grids, const double *d_k, const double *d_pk, double *d_ws,
                       int32_t *d_ranges)
{
    hipLaunchKernelGGL(k_resample, dim3(nchunk), dim3(64), 0, (hipStream_t)stream, d_chunks, d_spec, d_grids, d_k,
                       d_pk, d_ws, d_ranges);
    return (int)hipGetLastError();
}

static size_t fft_lds_bytes(int max_M)
{
    return ((size_t)lds_data_len(max_M) + 64 + (size_t)(max_M + 63) / 64) * sizeof(cd);
}

int ce_launch_prepare(void *stream, int nspec, const ce_spec_dev *d_spec, const ce_grid_dev *d_grids, int nq,
                      const double *d_ws, double *d_xf, int threads, int max_M)
{
    const size_t lds = fft_lds_bytes(max_M);
    int st = set_lds((const void *)k_prepare, lds);
    if (st) return st;
    hipLaunchKernelGGL(k_prepare, dim3(nspec), dim3(threads), lds, (hipStream_t)stream, d_spec, d_grids, nq, d_ws,
                       reinterpret_cast<cd *>(d_xf), max_M);
    return (int)hipGetLastError();
}

int ce_launch_tables(void *stream, int nunit, const ce_unit_dev *d_units, const ce_spec_dev *d_spec,
                     const ce_grid_dev *d_grids, const ce_table_desc *d_tabs, int ntab, const double *d_xf,
                     int nsep, const int32_t *d_sidx, const double *d_sep, double *d_itab, double *d_full,
                     int64_t full_stride, const double *d_A, const double *d_gout, int nz, int nl, double *d_xbar,
                     double *d_wbuf, int threads, int max_M, const int32_t *d_packs, int npack, int lt_off)
{
    table_args a;
    a.units = d_units;
    a.specs = d_spec;
    a.grids = d_grids;
    a.tabs = d_tabs;
    a.ntab = ntab;
    a.xf = reinterpret_cast<const cd *>(d_xf);
    a.nsep = nsep;
    a.sidx = d_sidx;
    a.sep = d_sep;
    a.itab = d_itab;
    a.full = d_full;
    a.full_stride = full_stride;
    a.A = d_A;
    a.gout = d_gout;
    a.nz = nz;
    a.nl = nl;
    a.xbar = reinterpret_cast<cd *>(d_xbar);
    a.wbuf = d_wbuf;
    a.max_M = max_M;
    a.packs = reinterpret_cast<const int4 *>(d_packs);
    a.lt_off = (d_packs && lt_off > 0) ? lt_off : lds_data_len(max_M);
    const size_t lds = ((size_t)a.lt_off + 64 + (size_t)(max_M + 63) / 64) * sizeof(cd);
    int st = set_lds((const void *)k_tables, lds);
    if (st) return st;
    hipLaunchKernelGGL(k_tables, dim3(d_packs ? npack : nunit), dim3(threads), lds, (hipStream_t)stream, a);
    return (int)hipGetLastError();
}

int ce_launch_project(void *stream, int nspec, int ntab, int nz, int nl, int nsep, const double *d_A,
                      const double *d_itab, double *d_out)
{
    const int64_t total = (int64_t)nspec * nz * nl * nsep;
    const int threads = 256;
    int64_t blocks = (total + threads - 1) / threads;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_project, dim3((unsigned)blocks), dim3(threads), 0, (hipStream_t)stream, nspec, ntab, nz,
                       nl, nsep, d_A, d_itab, d_out);
    return (int)hipGetLastError();
}

int ce_launch_adj_c2r(void *stream, int nspec, const ce_spec_dev *d_spec, const ce_grid_dev *d_grids,
                      const ce_table_desc *d_tabs, int ntab, int nq, double *d_ws, const double *d_xbar, int threads,
                      int max_M, const double *d_A, const double *d_itab, double *d_out, int nz, int nl, int nsep)
{
    const size_t lds = fft_lds_bytes(max_M);
    int st = set_lds((const void *)k_adj_c2r, lds);
    if (st) return st;
    proj_args pj;
    pj.A = d_A;
    pj.itab = d_itab;
    pj.out = (d_A && d_itab) ? d_out : nullptr;
    pj.nz = nz;
    pj.nl = nl;
    pj.nsep = nsep;
    hipLaunchKernelGGL(k_adj_c2r, dim3(nspec), dim3(threads), lds, (hipStream_t)stream, d_spec, d_grids, d_tabs, ntab,
                       nq, d_ws, reinterpret_cast<const cd *>(d_xbar), max_M, pj);
    return (int)hipGetLastError();
}

int ce_launch_spline_adj(void *stream, int nchunk, const ce_chunk_dev *d_chunks, const ce_spec_dev *d_spec,
                         const ce_grid_dev *d_grids, const double *d_k, const double *d_ws, double *d_gpk,
                         const int32_t *d_ranges)
{
    hipLaunchKernelGGL(k_spline_adj, dim3(nchunk), dim3(64), 0, (hipStream_t)stream, d_chunks, d_spec, d_grids, d_k,
                       d_ws, d_gpk, d_ranges);
    return (int)hipGetLastError();
}

}  // extern "C"
