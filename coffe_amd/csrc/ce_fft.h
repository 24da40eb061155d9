/*
 * LDS-resident complex FFT building blocks shared by the HIP sources (gfx950, FP64).
 *
 * One workgroup owns one transform.  The length-M complex sequence lives in LDS as
 * interleaved doubles and is transformed by self-sorting Stockham passes: every
 * thread first pulls the R inputs of each of its butterflies into registers, the
 * workgroup synchronises, then every thread applies twiddles, the radix-R DFT and
 * stores the R outputs.  A single LDS buffer of 16*M bytes is therefore enough
 * (M <= 8192 -> 128 KiB of the 160 KiB per CU).  Callers launch at least M/8 threads.
 */
#ifndef COFFE_ENGINE_FFT_H
#define COFFE_ENGINE_FFT_H

#include <hip/hip_runtime.h>

#include "ce_internal.h"

namespace ce {

// 16-byte aligned so that LDS and global accesses of a complex value are single
// 128-bit operations (ds_read_b128 / global_load_dwordx4)
struct __attribute__((aligned(16))) cd {
    double x, y;
};

__device__ __forceinline__ cd mk(double a, double b) { cd r; r.x = a; r.y = b; return r; }
__device__ __forceinline__ cd cadd(cd a, cd b) { return mk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cd csub(cd a, cd b) { return mk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ cd cmul(cd a, cd b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ cd cconj(cd a) { return mk(a.x, -a.y); }
__device__ __forceinline__ cd cscale(cd a, double s) { return mk(a.x * s, a.y * s); }
__device__ __forceinline__ cd mul_mi(cd a) { return mk(a.y, -a.x); }   // * (-i)
__device__ __forceinline__ cd mul_pi(cd a) { return mk(-a.y, a.x); }   // * (+i)

// ---- global-memory views ------------------------------------------------------
// Pointers read from memory (the fields of the grid and batch structures) are
// generic pointers, and a generic (flat) load might address LDS: the compiler keeps
// it behind every earlier LDS store and waits for it on both counters, which
// serialises load-compute-store loops.  Through these views the loads are global
// loads, which it issues early and waits for one at a time.
typedef __attribute__((address_space(1))) const double gdouble;
typedef __attribute__((address_space(1))) const cd gcd;
typedef __attribute__((address_space(1))) double gdouble_out;
typedef __attribute__((address_space(1))) cd gcd_out;
__device__ __forceinline__ gdouble *as_global(const double *p) { return (gdouble *)p; }
__device__ __forceinline__ gcd *as_global(const cd *p) { return (gcd *)p; }
__device__ __forceinline__ gdouble_out *as_global_out(double *p) { return (gdouble_out *)p; }
__device__ __forceinline__ gcd_out *as_global_out(cd *p) { return (gcd_out *)p; }
// a complex value through a global view (one 128-bit access: the type is 16-byte aligned)
__device__ __forceinline__ cd ldg(gcd *p) { return mk(p->x, p->y); }
__device__ __forceinline__ void stg(gcd_out *p, cd v) { p->x = v.x; p->y = v.y; }

// ---- workgroup barrier for LDS traffic ---------------------------------------------
// The transforms exchange data through LDS only.  A full __syncthreads() also waits
// for every global load in flight; this barrier orders the LDS accesses alone, so
// loads issued ahead of a transform (operands of the phase after it) stay in flight
// across all its passes.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ---- LDS layout ----------------------------------------------------------------
// Complex element i of a transform lives at LDS slot pidx(i) = i ^ ((i >> 3) & 7): the
// low three bits are XORed with the next three, a permutation inside each aligned block
// of eight slots.  gfx950 serves a ds_write_b128 eight consecutive lanes a cycle over
// 128 bytes of banks (slot mod 8) and a ds_read_b128 sixteen lanes a cycle over 256
// bytes (slot mod 16; lanes {0-3,12-15,20-23,24-27} and {4-7,8-11,16-19,28-31}, the
// same +32): the swizzle keeps the stride-R stores of the early Stockham passes and
// the unit-stride loads on distinct banks for power-of-two lengths (a pad slot every
// 16 elements, the layout this replaces, cost each load a second cycle).
__device__ __forceinline__ int pidx(int i) { return i ^ ((i >> 3) & 7); }
__host__ __device__ __forceinline__ int lds_data_len(int M) { return CE_LDS_DATA_LEN(M); }
// real element i of a transform stored as complex pairs (x_{2j}, x_{2j+1})
__device__ __forceinline__ double &lds_real(cd *s, int i)
{
    return reinterpret_cast<double *>(s)[2 * pidx(i >> 1) + (i & 1)];
}

// ---- radix-R forward DFT butterflies (sign -1) --------------------------------
template <int R> __device__ __forceinline__ void dft(cd *v);

template <> __device__ __forceinline__ void dft<2>(cd *v)
{
    cd a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <> __device__ __forceinline__ void dft<3>(cd *v)
{
    const double s3 = 0.86602540378443864676372317075294;
    cd a = v[0], s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    v[0] = cadd(a, s);
    cd m = mk(a.x - 0.5 * s.x, a.y - 0.5 * s.y);
    cd rot = mul_mi(cscale(d, s3));
    v[1] = cadd(m, rot);
    v[2] = csub(m, rot);
}

template <> __device__ __forceinline__ void dft<4>(cd *v)
{
    cd s0 = cadd(v[0], v[2]), d0 = csub(v[0], v[2]);
    cd s1 = cadd(v[1], v[3]), d1 = mul_mi(csub(v[1], v[3]));
    v[0] = cadd(s0, s1);
    v[2] = csub(s0, s1);
    v[1] = cadd(d0, d1);
    v[3] = csub(d0, d1);
}

template <> __device__ __forceinline__ void dft<5>(cd *v)
{
    const double c1 = 0.30901699437494742410229341718282;
    const double c2 = -0.80901699437494742410229341718282;
    const double s1 = 0.95105651629515357211643933337938;
    const double s2 = 0.58778525229247312916870595463907;
    cd a = v[0];
    cd b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]);
    cd e1 = csub(v[1], v[4]), e2 = csub(v[2], v[3]);
    v[0] = cadd(a, cadd(b1, b2));
    cd m1 = mk(a.x + c1 * b1.x + c2 * b2.x, a.y + c1 * b1.y + c2 * b2.y);
    cd m2 = mk(a.x + c2 * b1.x + c1 * b2.x, a.y + c2 * b1.y + c1 * b2.y);
    cd n1 = mul_mi(mk(s1 * e1.x + s2 * e2.x, s1 * e1.y + s2 * e2.y));
    cd n2 = mul_mi(mk(s2 * e1.x - s1 * e2.x, s2 * e1.y - s1 * e2.y));
    v[1] = cadd(m1, n1);
    v[4] = csub(m1, n1);
    v[2] = cadd(m2, n2);
    v[3] = csub(m2, n2);
}

template <> __device__ __forceinline__ void dft<8>(cd *v)
{
    const double r2 = 0.70710678118654752440084436210485;
    cd e[4] = {v[0], v[2], v[4], v[6]};
    cd o[4] = {v[1], v[3], v[5], v[7]};
    dft<4>(e);
    dft<4>(o);
    // o[k] *= W8^k
    cd o1 = mk(r2 * (o[1].x + o[1].y), r2 * (o[1].y - o[1].x));     // (1-i)/sqrt2
    cd o2 = mul_mi(o[2]);                                           // -i
    cd o3 = mk(r2 * (o[3].y - o[3].x), -r2 * (o[3].x + o[3].y));    // (-1-i)/sqrt2
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
}

// radix 10 as 2 x 5 in place: x[5 n1 + n2] -> sums (n1 = 0 + 1) in v[0..4] and twiddled
// differences in v[5..9], DFT5 of each half, X[2 k2] = v[k2] and X[2 k2 + 1] = v[5 + k2]
template <> __device__ __forceinline__ void dft<10>(cd *v)
{
    const double c1 = 0.80901699437494742410229341718282;   // cos(pi/5)
    const double s1 = 0.58778525229247312916870595463907;   // sin(pi/5)
    const double c2 = 0.30901699437494742410229341718282;   // cos(2 pi/5)
    const double s2 = 0.95105651629515357211643933337938;   // sin(2 pi/5)
#pragma unroll
    for (int n2 = 0; n2 < 5; ++n2) {
        const cd a = v[n2], b = v[5 + n2];
        v[n2] = cadd(a, b);
        v[5 + n2] = csub(a, b);
    }
    // W10^m = e^{-i pi m/5}
    v[6] = cmul(v[6], mk(c1, -s1));
    v[7] = cmul(v[7], mk(c2, -s2));
    v[8] = cmul(v[8], mk(-c2, -s2));
    v[9] = cmul(v[9], mk(-c1, -s1));
    dft<5>(v);
    dft<5>(v + 5);
    const cd t1 = v[1], t2 = v[2], t3 = v[3], t4 = v[4], t5 = v[5], t6 = v[6], t7 = v[7], t8 = v[8];
    v[2] = t1; v[4] = t2; v[6] = t3; v[8] = t4;
    v[1] = t5; v[3] = t6; v[5] = t7; v[7] = t8;
}

// Twiddles W_M^k = e^{-2 pi i k / M} come from two small LDS tables,
// W_M^k = T1[k >> 6] * T0[k & 63] (T0: 64 entries, T1: M/64 entries), both filled
// from the correctly rounded e^{-2 pi i k/N} table of the grid; the product adds at
// most one rounding.
__device__ __forceinline__ cd twiddle(const cd *lt, int k) { return cmul(lt[64 + (k >> 6)], lt[k & 63]); }

// fill the LDS twiddle tables of an M-point transform from tw (e^{-2 pi i k/(stride M)}):
// stride 2 for the half-length transform of an N = 2M grid table, 1 for a table of
// the transform's own length
__device__ __forceinline__ void load_twiddles(cd *lt, int M, gcd *tw, int stride = 2)
{
    const int nhi = (M + 63) >> 6;
    for (int i = threadIdx.x; i < 64 + nhi; i += blockDim.x) {
        const int k = (i < 64) ? i : 64 * (i - 64);
        lt[i] = (k < M) ? ldg(tw + stride * k) : mk(1.0, 0.0);
    }
}

// One Stockham pass with radix R over s[0..M); Ns = product of earlier radices and
// stride = M/(Ns R) (the twiddle step), both uniform over the wavefront; lt holds the
// LDS twiddle tables; threads tid < T of the group own the transform (every thread of
// the workgroup calls: the pass ends on workgroup barriers).
//
// Index arithmetic is kept off the per-butterfly path: j mod Ns is a mask when Ns is
// a power of two and otherwise a float-reciprocal quotient with one correction each
// way (j, Ns < 2^24, so the estimate is off by at most one); the swizzled LDS slots of
// a butterfly are one computed slot plus uniform offsets when the element stride (B
// for the loads, Ns for the stores) is a multiple of 64.
template <int R, int MAXB, bool TW>
__device__ __forceinline__ void fft_pass(cd *s, int Ns, int stride, const cd *__restrict__ lt, int tid, int T)
{
    if (!TW) Ns = 1;
    const int B = Ns * stride;   // M / R butterflies
    cd v[MAXB][R];
    if ((B & 63) == 0) {
        // pidx(j + r B) = pidx(j) + r B: a multiple of 64 leaves the six low bits alone
#pragma unroll
        for (int b = 0; b < MAXB; ++b) {
            const int j = tid + b * T;
            if (j < B) {
                const cd *p = s + pidx(j);
#pragma unroll
                for (int r = 0; r < R; ++r) v[b][r] = p[r * B];
            }
        }
    } else {
#pragma unroll
        for (int b = 0; b < MAXB; ++b) {
            const int j = tid + b * T;
            if (j < B) {
#pragma unroll
                for (int r = 0; r < R; ++r) v[b][r] = s[pidx(j + r * B)];
            }
        }
    }
    lds_barrier();
    const bool pow2 = (Ns & (Ns - 1)) == 0;
    const bool lin = (Ns & 63) == 0;   // pidx(base + r Ns) = pidx(base) + r Ns
    const float rns = __builtin_amdgcn_rcpf((float)Ns);
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int j = tid + b * T;
        if (j < B) {
            int jm;
            if (pow2) {
                jm = j & (Ns - 1);
            } else {
                jm = j - Ns * (int)((float)j * rns);
                jm += (jm < 0) ? Ns : 0;
                jm -= (jm >= Ns) ? Ns : 0;
            }
            if (TW) {
                // w_r = W^(r jm stride): one table lookup, the rest by binary splitting
                // w_r = w_{r/2} w_{r - r/2} (at most four roundings deep for R <= 10);
                // jm = 0 gives w_r = 1 exactly
                cd w[R];
                w[1] = twiddle(lt, jm * stride);
#pragma unroll
                for (int r = 2; r < R; ++r) w[r] = cmul(w[r >> 1], w[r - (r >> 1)]);
#pragma unroll
                for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], w[r]);
            }
            dft<R>(v[b]);
            const int base = (j - jm) * R + jm;
            if (lin) {
                cd *p = s + pidx(base);
#pragma unroll
                for (int r = 0; r < R; ++r) p[r * Ns] = v[b][r];
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) s[pidx(base + r * Ns)] = v[b][r];
            }
        }
    }
    lds_barrier();
}

// Butterflies per thread: a transform gets at least M/8 threads, so a radix-R pass
// has at most ceil(8/R) butterflies per thread; only those register shapes are
// instantiated (at most 12 complex values live per thread).
// TW = false for the first pass (Ns = 1: every twiddle is 1), a separate instance so
// that no register copies join the two paths.
template <int R, bool TW>
__device__ __forceinline__ void fft_pass_dispatch(cd *s, int Ns, int stride, const cd *__restrict__ lt, int tid, int T)
{
    const int B = Ns * stride;
    if (R >= 8 || B <= T) fft_pass<R, 1, TW>(s, Ns, stride, lt, tid, T);
    else if (R == 4 || R == 5 || B <= 2 * T) fft_pass<R, 2, TW>(s, Ns, stride, lt, tid, T);
    else fft_pass<R, 4, TW>(s, Ns, stride, lt, tid, T);
}

// one pass of radix code c (3 bits) at Ns; rem = M / Ns on entry, M / (Ns R) on exit
template <bool RADIX10, bool TW>
__device__ __forceinline__ void fft_pass_code(cd *s, uint32_t c, int &Ns, unsigned &rem, const cd *__restrict__ lt, int tid, int T)
{
    switch (c) {
    case 7:
        if (RADIX10) { rem /= 10; fft_pass_dispatch<10, TW>(s, Ns, (int)rem, lt, tid, T); Ns *= 10; }
        break;
    case 5: rem >>= 3; fft_pass_dispatch<8, TW>(s, Ns, (int)rem, lt, tid, T); Ns *= 8; break;
    case 3: rem >>= 2; fft_pass_dispatch<4, TW>(s, Ns, (int)rem, lt, tid, T); Ns *= 4; break;
    case 1: rem >>= 1; fft_pass_dispatch<2, TW>(s, Ns, (int)rem, lt, tid, T); Ns *= 2; break;
    case 4: rem /= 5; fft_pass_dispatch<5, TW>(s, Ns, (int)rem, lt, tid, T); Ns *= 5; break;
    case 2: rem /= 3; fft_pass_dispatch<3, TW>(s, Ns, (int)rem, lt, tid, T); Ns *= 3; break;
    default: break;
    }
}

// forward complex FFT of length M in LDS (natural order in and out) by the threads
// tid < T of a group (T >= M/8; every thread of the workgroup calls); M and code are
// uniform over each wavefront (a group is whole wavefronts); code = the packed radix
// schedule (ce_radix_plan.code, held in a scalar register: no memory access per pass);
// lt = LDS twiddle tables (load_twiddles), valid before the call
template <bool RADIX10>
__device__ __forceinline__ void fft_lds_t(cd *s, int M, uint32_t code, const cd *__restrict__ lt, int tid, int T)
{
    int Ns = 1;
    unsigned rem = (unsigned)__builtin_amdgcn_readfirstlane(M);   // M / Ns
    uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)code);
    if (c == 0) return;
    fft_pass_code<RADIX10, false>(s, c & 7, Ns, rem, lt, tid, T);
    for (c >>= 3; c != 0; c >>= 3) fft_pass_code<RADIX10, true>(s, c & 7, Ns, rem, lt, tid, T);
}
// the plans of ce_radix_schedule (radices 8, 4, 2, 5, 3)
__device__ __forceinline__ void fft_lds(cd *s, int M, uint32_t code, const cd *__restrict__ lt, int tid, int T)
{
    fft_lds_t<false>(s, M, code, lt, tid, T);
}
__device__ __forceinline__ void fft_lds(cd *s, int M, uint32_t code, const cd *__restrict__ lt)
{
    fft_lds_t<false>(s, M, code, lt, threadIdx.x, blockDim.x);
}
// the plans of ce_radix_schedule_ex with radix-10 passes (the kernels that instantiate
// the wider pass: its registers count toward every path of a kernel)
__device__ __forceinline__ void fft_lds10(cd *s, int M, uint32_t code, const cd *__restrict__ lt)
{
    fft_lds_t<true>(s, M, code, lt, threadIdx.x, blockDim.x);
}

// r2c split step: from Z = FFT_M(z), z_j = x_{2j} + i x_{2j+1}, produce X_k for the
// pair (k, M-k): X_k = (Z_k + conj Z_{M-k})/2 - i/2 W_N^k (Z_k - conj Z_{M-k}).
__device__ __forceinline__ cd r2c_split(cd Zk, cd Zmk, cd w)
{
    cd b = cconj(Zmk);
    cd e = cscale(cadd(Zk, b), 0.5);
    cd o = mul_mi(cscale(csub(Zk, b), 0.5));
    return cadd(e, cmul(w, o));
}

// c2r pre-step for the pair: Z_k = (X_k + conj X_{M-k}) + i e^{+2 pi i k/N}(X_k - conj X_{M-k})
__device__ __forceinline__ cd c2r_merge(cd Xk, cd Xmk, cd w)
{
    cd b = cconj(Xmk);
    cd e = cadd(Xk, b);
    cd o = cmul(csub(Xk, b), cconj(w));
    return cadd(e, mul_pi(o));
}

}  // namespace ce

#endif
