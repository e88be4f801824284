// qnd_device.hpp -- device code of libqnd for gfx950 (CDNA4, wave64).
//
// The split-operator step is built from "line passes": a workgroup loads a
// tile of TW lines of length n along one axis into LDS, runs an in-place
// Stockham FFT on them (radix 8/4/2, eight values per thread, twiddles from an
// FP64/FP32 table), applies the fused pointwise operators of the step,
// optionally transforms back, reduces observables in FP64 and writes the tile
// back.  Only the state crosses HBM; potential, CAP, kinetic, laser-coupling
// and mask factors are analytic or per-axis tables of n_j values.
//
// Pass kinds (one small kernel each, so the instruction stream of a pass
// holds only its own operators):
//   PK_PLAIN : one transform along an axis, times a scale
//   PK_KIN   : FFT -> kinetic factor (all axes in k) -> IFFT    (last axis)
//   PK_POT   : IFFT -> r-space operators -> FFT                  (axis 0)
//   PK_POT1D : PK_POT with the kinetic factors before/after (dim = 1)
// All grid extents are powers of two >= 8, so every index is a shift/mask.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace qnd {

// Workgroup barrier that no memory access is moved across.  With a bare
// __syncthreads() the compiler has sunk an LDS read of a Stockham stage
// below the barrier that separates the stage's reads from its writes (the
// read is first used after the twiddle block), so another wavefront's write
// could land first.  The empty asm statements with a memory clobber pin
// every load and store on its side of the barrier.
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("" ::: "memory");
    __syncthreads();
    asm volatile("" ::: "memory");
}

// Exchange fence for a line held by exactly one wavefront: the LDS operations
// of one wavefront complete in program order, so no workgroup barrier is
// needed; the fence keeps the compiler from moving LDS accesses across it.
__device__ __forceinline__ void wave_fence()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
}

// ---------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------
template <typename R> struct CT;
template <> struct CT<float> { using T = float2; };
template <> struct CT<double> { using T = double2; };
template <typename R> using C = typename CT<R>::T;

template <typename R> __device__ __forceinline__ C<R> mk(R x, R y) { C<R> z; z.x = x; z.y = y; return z; }
template <typename R> __device__ __forceinline__ C<R> add(C<R> a, C<R> b) { return mk<R>(a.x + b.x, a.y + b.y); }
template <typename R> __device__ __forceinline__ C<R> sub(C<R> a, C<R> b) { return mk<R>(a.x - b.x, a.y - b.y); }
template <typename R> __device__ __forceinline__ C<R> mul(C<R> a, C<R> b)
{
    return mk<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename R> __device__ __forceinline__ C<R> scl(C<R> a, R s) { return mk<R>(a.x * s, a.y * s); }
// multiply by (DIR * i): DIR=-1 -> -i z, DIR=+1 -> +i z
template <typename R, int DIR> __device__ __forceinline__ C<R> mul_di(C<R> z)
{
    return DIR < 0 ? mk<R>(z.y, -z.x) : mk<R>(-z.y, z.x);
}

// ---------------------------------------------------------------------
// in-register DFT kernels (natural order in and out), sign DIR
// ---------------------------------------------------------------------
template <typename R, int DIR> __device__ __forceinline__ void dft2(C<R> *x)
{
    C<R> a = x[0], b = x[1];
    x[0] = add<R>(a, b);
    x[1] = sub<R>(a, b);
}

template <typename R, int DIR> __device__ __forceinline__ void dft4(C<R> *x)
{
    C<R> a0 = add<R>(x[0], x[2]), a1 = sub<R>(x[0], x[2]);
    C<R> b0 = add<R>(x[1], x[3]), b1 = mul_di<R, DIR>(sub<R>(x[1], x[3]));
    x[0] = add<R>(a0, b0);
    x[2] = sub<R>(a0, b0);
    x[1] = add<R>(a1, b1);
    x[3] = sub<R>(a1, b1);
}

// t * (1/sqrt 2).  In FP32 the constant is a hi+lo pair: a single rounded
// 1/sqrt2 has 2 h^2 = 1 - 3.4e-8, a bias that every radix-8 stage of every
// step would apply to the norm.  The exact product t*hi and the small t*lo
// enter one fused rounding (fma(t, hi, t*lo)); rounding t*hi first would
// discard the correction (it is below half an ulp of the result).
template <typename R> __device__ __forceinline__ R mul_rsqrt2(R t);
template <> __device__ __forceinline__ double mul_rsqrt2<double>(double t)
{
    return t * 0.70710678118654752440084436210484903928;
}
template <> __device__ __forceinline__ float mul_rsqrt2<float>(float t)
{
    const float hi = 0.70710677f, lo = 1.2101617e-08f;  // hi + lo = 1/sqrt(2) to ~1e-15
    return fmaf(t, hi, t * lo);
}

// radix 8 as two radix-4 halves (even/odd inputs) and the w_8^q combine
template <typename R, int DIR> __device__ __forceinline__ void dft8(C<R> *x)
{
    C<R> e[4] = {x[0], x[2], x[4], x[6]};
    C<R> o[4] = {x[1], x[3], x[5], x[7]};
    dft4<R, DIR>(e);
    dft4<R, DIR>(o);
    C<R> o1 = mk<R>(mul_rsqrt2<R>(o[1].x - (R)DIR * o[1].y), mul_rsqrt2<R>(o[1].y + (R)DIR * o[1].x));
    C<R> o2 = mul_di<R, DIR>(o[2]);
    C<R> o3 = mk<R>(mul_rsqrt2<R>(-o[3].x - (R)DIR * o[3].y), mul_rsqrt2<R>(-o[3].y + (R)DIR * o[3].x));
    x[0] = add<R>(e[0], o[0]);
    x[4] = sub<R>(e[0], o[0]);
    x[1] = add<R>(e[1], o1);
    x[5] = sub<R>(e[1], o1);
    x[2] = add<R>(e[2], o2);
    x[6] = sub<R>(e[2], o2);
    x[3] = add<R>(e[3], o3);
    x[7] = sub<R>(e[3], o3);
}

// t * (sqrt(3)/2), FP32 as a hi+lo pair for the same reason as mul_rsqrt2
template <typename R> __device__ __forceinline__ R mul_sqrt3h(R t);
template <> __device__ __forceinline__ double mul_sqrt3h<double>(double t)
{
    return t * 0.86602540378443864676372317075293618347;
}
template <> __device__ __forceinline__ float mul_sqrt3h<float>(float t)
{
    const float hi = 0.8660254f, lo = 2.3801166e-08f;  // hi + lo = sqrt(3)/2 to ~1e-15
    return fmaf(t, hi, t * lo);
}

// radix 3: y_q = sum_r x_r w^{qr}, w = e^{DIR 2 pi i/3} = -1/2 + DIR i sqrt(3)/2
template <typename R, int DIR> __device__ __forceinline__ void dft3(C<R> *x)
{
    const C<R> s = add<R>(x[1], x[2]), d = sub<R>(x[1], x[2]);
    const C<R> m = mk<R>(x[0].x - (R)0.5 * s.x, x[0].y - (R)0.5 * s.y);
    const C<R> r = mul_di<R, DIR>(mk<R>(mul_sqrt3h<R>(d.x), mul_sqrt3h<R>(d.y)));
    x[0] = add<R>(x[0], s);
    x[1] = add<R>(m, r);
    x[2] = sub<R>(m, r);
}

template <typename R, int RAD, int DIR> __device__ __forceinline__ void dftR(C<R> *x)
{
    if (RAD == 2) dft2<R, DIR>(x);
    else if (RAD == 3) dft3<R, DIR>(x);
    else if (RAD == 4) dft4<R, DIR>(x);
    else dft8<R, DIR>(x);
}

// Twiddle table entries W_n^m = exp(-2 pi i m/n): double2 for complex128;
// for complex64 a float4 (cos_hi, sin_hi, cos_lo, sin_lo) so that the
// rounded table does not bias the modulus of the transform.
template <typename R> struct TwT;
template <> struct TwT<double> { using T = double2; };
template <> struct TwT<float> { using T = float4; };

template <typename R, int DIR> __device__ __forceinline__ C<R> twmul(C<R> x, typename TwT<R>::T w);
template <> __device__ __forceinline__ C<double> twmul<double, -1>(C<double> x, double2 w) { return mul<double>(x, w); }
template <> __device__ __forceinline__ C<double> twmul<double, +1>(C<double> x, double2 w)
{
    w.y = -w.y;
    return mul<double>(x, w);
}
template <int DIR> __device__ __forceinline__ C<float> twmul_f(C<float> x, float4 w)
{
    // each exact hi product enters an fma; the lo terms are added before the
    // roundings they correct.  The order of the hi products is fixed: every
    // twiddle angle is at least 2 pi / 4096 from 0 and from +-pi/2, so cos and
    // sin keep their second-order terms in hi (unlike the small phase factors
    // of fac_mul); choosing the order per twiddle measured 6.5 % slower on the
    // 512^2 step for no significant change of its norm drift
    const float c = w.x, s = DIR < 0 ? w.y : -w.y, cl = w.z, sl = DIR < 0 ? w.w : -w.w;
    const float re = fmaf(x.x, c, fmaf(-x.y, s, x.x * cl - x.y * sl));
    const float im = fmaf(x.x, s, fmaf(x.y, c, x.x * sl + x.y * cl));
    return mk<float>(re, im);
}
template <> __device__ __forceinline__ C<float> twmul<float, -1>(C<float> x, float4 w) { return twmul_f<-1>(x, w); }
template <> __device__ __forceinline__ C<float> twmul<float, +1>(C<float> x, float4 w) { return twmul_f<+1>(x, w); }

// A twiddle or factor table entry widened to the arithmetic type CR: the
// passes that keep complex64 states in FP64 registers (the complex64 storage,
// FP64 arithmetic form) read the complex64 tables, whose FP32 hi + lo pairs
// hold the FP64 values to about 1e-15.
template <typename CR, typename TE> struct TwCvt {
    static __device__ __forceinline__ typename TwT<CR>::T f(const TE &e) { return e; }
};
template <> struct TwCvt<double, float4> {
    static __device__ __forceinline__ double2 f(const float4 &e)
    {
        return make_double2((double)e.x + (double)e.z, (double)e.y + (double)e.w);
    }
};

// Phase factor e^{arg}(cos ph - i sin ph) applied to a value: complex128
// multiplies by the FP64 factor; complex64 carries the FP64 factor as an
// FP32 hi+lo pair (compensated product), which keeps the propagation
// unitary to FP32 rounding instead of drifting by a per-angle bias.
template <typename R> struct Fac;
template <> struct Fac<double> { double2 v; };
// complex64: the factor is (-i)^q (hi + lo), hi + lo an FP32 pair with
// |hi.x| >= |hi.y| (the reduced angle, or a rotated table entry)
template <> struct Fac<float> { float2 hi, lo; int q; };
template <typename R> __device__ __forceinline__ C<R> fac_mul(C<R> x, const Fac<R> &f);
template <> __device__ __forceinline__ C<double> fac_mul<double>(C<double> x, const Fac<double> &f)
{
    return mul<double>(x, f.v);
}
// x is first rotated by (-i)^q, exactly (a swap and signs); then each
// component adds its hi.x product last.  Since |hi.x| >= |hi.y|, the inner
// fma rounds at the smaller product's scale and the lo terms reach the final
// rounding.  With a factor near 1 (cos θ rounds to 1.0f below θ ~ 2.4e-4 and
// the -θ²/2 sits in lo) the opposite order rounded x.y·hi.x + lo first and
// dropped the correction every time: a shear that grew the norm by about
// θ²/2 per application (1e-9 to 5e-9 for θ in [1e-5, 1e-3], float
// emulation).  Small angles are common: kinetic factors near k = 0, weak
// potentials.
template <> __device__ __forceinline__ C<float> fac_mul<float>(C<float> x, const Fac<float> &f)
{
    // (-i)^q x: q = 0 (a, b), 1 (b, -a), 2 (-a, -b), 3 (-b, a)
    const bool odd = (f.q & 1) != 0;
    const float u = odd ? x.y : x.x, v = odd ? x.x : x.y;
    const float xr = (f.q == 2 || f.q == 3) ? -u : u, xi = (f.q == 1 || f.q == 2) ? -v : v;
    const float lr = xr * f.lo.x - xi * f.lo.y, li = xr * f.lo.y + xi * f.lo.x;
    return mk<float>(fmaf(xr, f.hi.x, fmaf(-xi, f.hi.y, lr)), fmaf(xi, f.hi.x, fmaf(xr, f.hi.y, li)));
}

// ---------------------------------------------------------------------
// Register-resident Stockham transform of one line.
//
// T = n / VPT threads share a line; thread t holds x[m] = element t + m*T
// (m < VPT) when the transform starts and again when it ends, so the global
// load, the pointwise operators and the global store all work on registers.
// Stage (radix RAD, span Ns) of the Stockham autosort: the thread owns the
// butterflies j = t + b*T (b < VPT/RAD);
//   inputs  y[r] = elem[j + r*n/RAD]
//   twiddle y[r] *= W_n^{r k n/(Ns RAD)},  k = j mod Ns
//   outputs elem[(j-k)*RAD + k + r*Ns] = DFT_RAD(y)[r]
// The first stage's inputs (Ns = 1) are t + T*(b + r*VPT/RAD) and the last
// stage's outputs (Ns = n/RAD, so k = j) are the same set: both map onto
// x[b + r*VPT/RAD] without LDS.  Interior stages exchange through two LDS
// buffers used alternately (one barrier per exchange: a buffer is written
// again only after every thread has passed the barrier that follows its
// reads).  `wc` counts the exchanges of the kernel and picks the buffer.
// ---------------------------------------------------------------------
// Butterflies of one stage; twf(b, r) gives the twiddle of input r (>= 1) of
// butterfly b (not called when Ns = 1).  IN: inputs from the LDS buffer lin (else from registers, the
// first stage); OUT: outputs to the LDS buffer lout (else to registers, the
// last stage).  Both are compile-time, so the element moves carry no tests.
template <typename R, int RAD, int DIR, int VPT, bool IN, bool OUT, class TWF>
__device__ __forceinline__ void stage_core(C<R> (&x)[VPT], const C<R> *lin, C<R> *lout, int n, int T, int t,
                                           int Ns, const TWF &twf)
{
    if constexpr (VPT % RAD == 0) {
        constexpr int NB = VPT / RAD;
        const int nR = n / RAD;
        typename TwT<R>::T w[VPT];
        if (Ns > 1) {  // twiddles first: a table's load latency overlaps the LDS reads
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int r = 1; r < RAD; ++r) w[b * RAD + r] = twf(b, r);
        }
        C<R> y[VPT];
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int r = 0; r < RAD; ++r) {
                if constexpr (IN) y[b * RAD + r] = lin[t + b * T + r * nR];
                else y[b * RAD + r] = x[b + r * NB];
            }
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            C<R> *yb = y + b * RAD;
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < RAD; ++r) yb[r] = twmul<R, DIR>(yb[r], w[b * RAD + r]);
            }
            dftR<R, RAD, DIR>(yb);
        }
        if constexpr (OUT) {
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int j = t + b * T, k = j & (Ns - 1);
                const int o = (j - k) * RAD + k;
#pragma unroll
                for (int r = 0; r < RAD; ++r) lout[o + r * Ns] = y[b * RAD + r];
            }
        } else {
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int r = 0; r < RAD; ++r) x[b + r * NB] = y[b * RAD + r];
        }
    }
}

// Twiddles of four-value-per-thread lines (n = 2^log2n <= 2048, T = n/4):
// every twiddle a thread needs is read once, at kernel entry, into registers.
// Both transforms of a pass share them (the direction only conjugates), so
// no stage waits on a table load: these grids run one wavefront per SIMD,
// where each such wait is exposed.
#define QND_TW4_STAGES 4   // radix-4 stages with twiddles: s = 1..4 (n <= 2048)
template <typename R> struct Tw4 {
    typename TwT<R>::T w4[QND_TW4_STAGES][3];  // stage s: W^{r k n/(4 Ns)}, k = t mod Ns, r = 1..3
    typename TwT<R>::T w2[2];                   // last radix-2 stage (odd log2 n): W^{t + b T}
};
template <typename R, typename TE>
__device__ __forceinline__ void tw4_load(Tw4<R> &P, const TE *__restrict__ tw, int log2n, int T, int t)
{
    const int ns4 = log2n >> 1;
#pragma unroll
    for (int s = 1; s <= QND_TW4_STAGES; ++s) {
        if (s < ns4) {
            const int k = t & ((1 << (2 * s)) - 1);
            const int stride = 1 << (log2n - 2 - 2 * s);
#pragma unroll
            for (int r = 1; r < 4; ++r) P.w4[s - 1][r - 1] = TwCvt<R, TE>::f(tw[r * k * stride]);
        }
    }
    if (log2n & 1) {
        P.w2[0] = TwCvt<R, TE>::f(tw[t]);
        P.w2[1] = TwCvt<R, TE>::f(tw[t + T]);
    }
}

// Stage schedule of a line of n = 2^L points (VPT 4 or 8) or n = 3 * 2^L
// points (VPT 12), fixed at compile time:
//   VPT 4:  radix-4 stages, then radix 2 if L is odd;
//   VPT 8:  radix-8 stages, then radix 4 or 2 for the remainder;
//   VPT 12: radix-4 stages, radix 2 if L is odd, radix 3 last (every span Ns
//           before it is then a power of two, so k = j mod Ns is a mask).
template <int VPT, int L> struct Sched {
    static constexpr int MAINR = (VPT == 8) ? 8 : 4;
    static constexpr int NMAIN = (VPT == 8) ? L / 3 : L / 2;
    static constexpr int REM = (VPT == 8) ? L % 3 : (L & 1);
    static constexpr int NST = NMAIN + (REM != 0 ? 1 : 0) + (VPT == 12 ? 1 : 0);
    static constexpr int N = (VPT == 12 ? 3 : 1) << L;
    static constexpr int rad(int s)
    {
        return s < NMAIN ? MAINR : (s == NMAIN && REM != 0) ? ((VPT == 8 && REM == 2) ? 4 : 2) : 3;
    }
    static constexpr int span(int s) { return s == 0 ? 1 : span(s - 1) * rad(s - 1); }
};
// Stage S of the schedule, then the rest (compile-time recursion).  The
// first stage reads registers, the last writes them, the others exchange
// through the two LDS buffers: stage S of a transform that starts at the
// kernel's exchange count wc0 writes buffer (wc0 + S) mod 2.  Twiddles come
// from the registers of Tw4 (VPT 4) or from the table tw.
template <typename R, int DIR, int VPT, int L, bool WL, int S, typename TE>
struct FftStages {
    static __device__ __forceinline__ void run(C<R> (&x)[VPT], C<R> *bufA, C<R> *bufB, int wc0, int t,
                                               const TE *__restrict__ tw, const Tw4<R> &P)
    {
        using Sc = Sched<VPT, L>;
        if constexpr (S < Sc::NST) {
            constexpr int n = Sc::N, T = n / VPT, RAD = Sc::rad(S), Ns = Sc::span(S);
            constexpr bool IN = S > 0, OUT = S < Sc::NST - 1;
            const C<R> *lin = ((wc0 + S - 1) & 1) ? bufB : bufA;
            C<R> *lout = ((wc0 + S) & 1) ? bufB : bufA;
            if constexpr (VPT == 4 && RAD == 4) {
                constexpr int SW = S > 0 ? S - 1 : 0;
                auto twf = [&](int, int r) { return P.w4[SW][r - 1]; };
                stage_core<R, RAD, DIR, VPT, IN, OUT>(x, lin, lout, n, T, t, Ns, twf);
            } else if constexpr (VPT == 4) {
                auto twf = [&](int b, int) { return P.w2[b]; };
                stage_core<R, RAD, DIR, VPT, IN, OUT>(x, lin, lout, n, T, t, Ns, twf);
            } else {
                constexpr int stride = n / (Ns * RAD);
                auto twf = [&](int b, int r) { return TwCvt<R, TE>::f(tw[r * ((t + b * T) & (Ns - 1)) * stride]); };
                stage_core<R, RAD, DIR, VPT, IN, OUT>(x, lin, lout, n, T, t, Ns, twf);
            }
            if constexpr (OUT) {
                if constexpr (WL) wave_fence();
                else lds_barrier();
            }
            FftStages<R, DIR, VPT, L, WL, S + 1, TE>::run(x, bufA, bufB, wc0, t, tw, P);
        }
    }
};
template <typename R, int DIR, int VPT, int L, typename TE>
__device__ __forceinline__ void fft_fixed(C<R> (&x)[VPT], C<R> *bufA, C<R> *bufB, int &wc, int t,
                                          const TE *__restrict__ tw, const Tw4<R> &P, bool wave_line)
{
    // wavefront-local exchanges exist only for 64-thread lines
    if constexpr (Sched<VPT, L>::N / VPT == 64) {
        if (wave_line) FftStages<R, DIR, VPT, L, true, 0, TE>::run(x, bufA, bufB, wc, t, tw, P);
        else FftStages<R, DIR, VPT, L, false, 0, TE>::run(x, bufA, bufB, wc, t, tw, P);
    } else {
        FftStages<R, DIR, VPT, L, false, 0, TE>::run(x, bufA, bufB, wc, t, tw, P);
    }
    wc += Sched<VPT, L>::NST - 1;  // exchanges of the transform
}

// Full transform of the line whose two LDS buffers are bufA / bufB: one
// uniform switch on the line length picks its fixed schedule.  log2n is
// log2 n (VPT 4, 8; n <= 2048 resp. 4096) or log2(n/3) (VPT 12, n <= 3072).
// Every thread of the block calls it (barriers inside).  TE: the twiddle
// table's entry type (a complex64 table for FP64 arithmetic is widened).
template <typename R, int DIR, int VPT, typename TE>
__device__ __forceinline__ void transform(C<R> (&x)[VPT], C<R> *bufA, C<R> *bufB, int &wc, int n, int log2n, int T,
                                          int t, const TE *__restrict__ tw, const Tw4<R> &P, bool wave_line)
{
#define QND_FFT_CASE(Lc)                                                                                            \
    case Lc: fft_fixed<R, DIR, VPT, Lc, TE>(x, bufA, bufB, wc, t, tw, P, wave_line); break;
    if constexpr (VPT == 4) {
        switch (log2n) {
            QND_FFT_CASE(3) QND_FFT_CASE(4) QND_FFT_CASE(5) QND_FFT_CASE(6) QND_FFT_CASE(7) QND_FFT_CASE(8)
            QND_FFT_CASE(9) QND_FFT_CASE(10) default: fft_fixed<R, DIR, VPT, 11, TE>(x, bufA, bufB, wc, t, tw, P,
                                                                                      wave_line);
        }
    } else if constexpr (VPT == 8) {
        switch (log2n) {
            QND_FFT_CASE(3) QND_FFT_CASE(4) QND_FFT_CASE(5) QND_FFT_CASE(6) QND_FFT_CASE(7) QND_FFT_CASE(8)
            QND_FFT_CASE(9) QND_FFT_CASE(10) QND_FFT_CASE(11) default: fft_fixed<R, DIR, VPT, 12, TE>(x, bufA, bufB,
                                                                                                       wc, t, tw, P,
                                                                                                       wave_line);
        }
    } else {
        switch (log2n) {
            QND_FFT_CASE(2) QND_FFT_CASE(3) QND_FFT_CASE(4) QND_FFT_CASE(5) QND_FFT_CASE(6) QND_FFT_CASE(7)
            QND_FFT_CASE(8) QND_FFT_CASE(9) default: fft_fixed<R, DIR, VPT, 10, TE>(x, bufA, bufB, wc, t, tw, P,
                                                                                     wave_line);
        }
    }
#undef QND_FFT_CASE
}

// ---------------------------------------------------------------------
// physics parameters (by value in every launch)
// ---------------------------------------------------------------------
#define QND_MAXD 4
struct Phys {
    int dim;
    int n[QND_MAXD];
    int lg[QND_MAXD];        // log2 n (when every extent is a power of two)
    int pow2;                // every extent a power of two: shift/mask indexing
    int off[QND_MAXD];       // offsets of each axis in the per-axis tables
    double dr[QND_MAXD];
    double L[QND_MAXD];      // n*dr
    double dp[QND_MAXD];     // 2 pi / L
    int pot_kind;            // 0 soft-core, 1 gaussian (Ne_short_2D)
    double pot_b, pot_c2;
    int cap_on, mask_on;
    const double *capt;      // per-axis CAP contributions cap0 (|x|-L_c)^2 or 0
    const double *maskt;     // per-axis cos^(1/8) mask factors
    double dNr;
};

__device__ __forceinline__ double coord_x(const Phys &P, int j, int i) { return -.5 * P.L[j] + i * P.dr[j]; }
__device__ __forceinline__ double coord_p(const Phys &P, int j, int i)
{
    return (i < (P.n[j] >> 1)) ? i * P.dp[j] : (i - P.n[j]) * P.dp[j];
}
// V(r) and dV/du (u = r^2) for the named potentials (src/meshgrid.c:140-153),
// in the precision of the state: complex128 evaluates them in FP64 like the
// reference; complex64 in FP32 (a relative change of V of a few 1e-7, below
// the FP32 state rounding).
// One point of a potential family known at compile time: SOFT (soft-core,
// PK = 0) with or without the Gaussian term (HASB), else the Gaussian well.
template <typename R, bool SOFT, bool HASB> struct PotPoint;
template <bool SOFT, bool HASB> struct PotPoint<double, SOFT, HASB> {
    static __device__ __forceinline__ double eval(const Phys &P, double rr, double &dVdu)
    {
        double rad = sqrt(rr);
        double u = rad * rad;
        if constexpr (SOFT) {
            double den = sqrt(u + P.pot_c2);
            double g = HASB ? P.pot_b * exp(-u) : 0.0;
            dVdu = g / den + .5 * (1. + g) / (den * (u + P.pot_c2));
            return -(1. + g) / den;
        } else {
            double g = 2.7459 * exp(-.5 * rad * rad / 1.96);
            dVdu = g * .5 / 1.96;
            return -g;
        }
    }
};
template <bool SOFT, bool HASB> struct PotPoint<float, SOFT, HASB> {
    static __device__ __forceinline__ float eval(const Phys &P, double rr, float &dVdu)
    {
        const float u = (float)rr;
        if constexpr (SOFT) {
            const float id = rsqrtf(u + (float)P.pot_c2);
            const float g = HASB ? (float)P.pot_b * __expf(-u) : 0.0f;
            dVdu = g * id + .5f * (1.f + g) * (id * id * id);
            return -(1.f + g) * id;
        } else {
            const float g = 2.7459f * __expf(u * (-.5f / 1.96f));
            dVdu = g * (.5f / 1.96f);
            return -g;
        }
    }
};
// FP64 potential for the complex64 passes (QND_POT_F64): 1/sqrt(u + c2)
// from an FP32 seed refined by two FP64 Newton steps (relative error ~1e-14,
// well below the FP32 state's rounding) instead of FP64 sqrt and divide;
// the exponentials in FP64.
template <bool SOFT, bool HASB> struct PotPointMix {
    static __device__ __forceinline__ double eval(const Phys &P, double rr, double &dVdu)
    {
        if constexpr (SOFT) {
            const double w = rr + P.pot_c2;
            double y = (double)rsqrtf((float)w);
            y = y * (1.5 - 0.5 * w * y * y);
            y = y * (1.5 - 0.5 * w * y * y);
            const double g = HASB ? P.pot_b * exp(-rr) : 0.0;
            dVdu = g * y + .5 * (1. + g) * (y * y * y);
            return -(1. + g) * y;
        } else {
            const double g = 2.7459 * exp(-.5 * rr / 1.96);
            dVdu = g * .5 / 1.96;
            return -g;
        }
    }
};
template <typename R> __device__ __forceinline__ R potential(const Phys &P, double rr, R &dVdu)
{
    if (P.pot_kind != 0) return PotPoint<R, false, false>::eval(P, rr, dVdu);
    if (P.pot_b != 0.0) return PotPoint<R, true, true>::eval(P, rr, dVdu);
    return PotPoint<R, true, false>::eval(P, rr, dVdu);
}
// V and dV/du at the VPT points of a thread, rr[m] = r^2: the family is
// chosen once, outside the loop, so the points form independent chains.
// MIX: FP64 values for a complex64 pass (PotPointMix), else the state's
// precision (PotPoint).
template <typename R, int VPT, bool MIX = false>
__device__ __forceinline__ void potential_pts(const Phys &P, const double (&rr)[VPT], R (&V)[VPT], R (&dVdu)[VPT])
{
    if constexpr (MIX) {
        if (P.pot_kind != 0) {
#pragma unroll
            for (int m = 0; m < VPT; ++m) V[m] = PotPointMix<false, false>::eval(P, rr[m], dVdu[m]);
        } else if (P.pot_b != 0.0) {
#pragma unroll
            for (int m = 0; m < VPT; ++m) V[m] = PotPointMix<true, true>::eval(P, rr[m], dVdu[m]);
        } else {
#pragma unroll
            for (int m = 0; m < VPT; ++m) V[m] = PotPointMix<true, false>::eval(P, rr[m], dVdu[m]);
        }
    } else if (P.pot_kind != 0) {
#pragma unroll
        for (int m = 0; m < VPT; ++m) V[m] = PotPoint<R, false, false>::eval(P, rr[m], dVdu[m]);
    } else if (P.pot_b != 0.0) {
#pragma unroll
        for (int m = 0; m < VPT; ++m) V[m] = PotPoint<R, true, true>::eval(P, rr[m], dVdu[m]);
    } else {
#pragma unroll
        for (int m = 0; m < VPT; ++m) V[m] = PotPoint<R, true, false>::eval(P, rr[m], dVdu[m]);
    }
}

// Derivatives with respect to the soft-core length a (c2 = a^2) of V and of
// dV/du, for the soft-core family V = -(1 + b e^{-u}) / sqrt(u + c2): the
// adjoint's parameter gradient.  Zero for the Gaussian well.
__device__ __forceinline__ void softcore_derivs(const Phys &P, double u, double &dV_da, double &dVdu_da)
{
    dV_da = 0.0;
    dVdu_da = 0.0;
    if (P.pot_kind != 0) return;
    const double a = sqrt(P.pot_c2), w = u + P.pot_c2;
    const double g = (P.pot_b != 0.0) ? P.pot_b * exp(-u) : 0.0;
    const double i1 = 1.0 / sqrt(w), i3 = i1 / w, i5 = i3 / w;
    dV_da = 2.0 * a * (.5 * (1.0 + g) * i3);
    dVdu_da = 2.0 * a * (-.5 * g * i3 - .75 * (1.0 + g) * i5);
}

#define QND_MAX_EXP_ARG 700.0     // src/integrators.h:12
#define QND_MAX_EXP_ARG_F32 80.0  // FP32 path: e^80 stays finite in float

template <typename R> struct Prec;
template <> struct Prec<double> { static constexpr double clamp = QND_MAX_EXP_ARG; };
template <> struct Prec<float> { static constexpr double clamp = QND_MAX_EXP_ARG_F32; };

// e^{arg} (cos ph - i sin ph), the angle reduced mod 2 pi in FP64 and the
// factor formed in FP64; complex64 receives it as an FP32 hi+lo pair.
template <typename R> __device__ Fac<R> phase_factor(double arg, double ph);
template <> __device__ __forceinline__ Fac<double> phase_factor<double>(double arg, double ph)
{
    double s, c;
    sincos(ph, &s, &c);
    // computed for every point and selected (no branch): e^0 = 1 exactly
    const double ex = exp(arg > QND_MAX_EXP_ARG ? QND_MAX_EXP_ARG : arg);
    const double e = (arg == 0.0) ? 1.0 : ex;
    Fac<double> f;
    f.v = mk<double>(e * c, -e * s);
    return f;
}
// sin r and cos r for |r| <= pi/4 (+ rounding), FP32 Taylor polynomials to
// r^9 / r^10: truncation below 3e-8, i.e. within the FP32 rounding.
__device__ __forceinline__ void sincos_poly_f(float r, float &s, float &c)
{
    const float r2 = r * r;
    float ps = fmaf(r2, 2.75573192e-6f, -1.98412698e-4f);   //  1/9!, -1/7!
    ps = fmaf(r2, ps, 8.33333333e-3f);                      //  1/5!
    ps = fmaf(r2, ps, -1.66666667e-1f);                     // -1/3!
    s = fmaf(r * r2, ps, r);
    float pc = fmaf(r2, -2.75573192e-7f, 2.48015873e-5f);   // -1/10!, 1/8!
    pc = fmaf(r2, pc, -1.38888889e-3f);                     // -1/6!
    pc = fmaf(r2, pc, 4.16666667e-2f);                      //  1/4!
    pc = fmaf(r2, pc, -0.5f);
    c = fmaf(r2, pc, 1.0f);
}
// 2^y for an FP64 y: integer part exact, fraction through the FP32 exp2
__device__ __forceinline__ float exp2_split(double y)
{
    const double n = floor(y);
    return ldexpf(exp2f((float)(y - n)), (int)n);
}
__device__ __forceinline__ float exp_f32(double arg)
{
    return exp2_split(arg * 1.4426950408889634073599246810019);
}
// complex64: the angle is reduced in FP64 to an octant (|r| <= pi/4), sin
// and cos come from FP32 polynomials, and the FP32 pair is corrected on its
// modulus: e = c^2 + s^2 - 1 is formed exactly (TwoProduct of both squares,
// Fast2Sum of c^2 - 1, Sterbenz for the rest) and lo = -e/2 * hi makes
// |hi + lo| = 1 to ~1e-14.  A residual error of the angle (~1e-7 rad) acts
// like a relative change of the Hamiltonian of that size; an error of the
// modulus would instead change the norm at every step.
//
// fac_octant_f32 forms that factor from the reduced angle rd, its quadrant q
// and the damping g; the two entry points below differ in how the angle is
// formed and reduced.
__device__ __forceinline__ Fac<float> fac_octant_f32(float rd, int q, float g)
{
#pragma clang fp contract(off)  // the error-free products and sums below must not be fused
    float sr, cr;
    sincos_poly_f(rd, sr, cr);
    const float pc = cr * cr, ec = fmaf(cr, cr, -pc);
    const float ps = sr * sr, es = fmaf(sr, sr, -ps);
    const float t = pc - 1.0f, terr = pc - (t + 1.0f);
    const float e = ((t + ps) + terr) + (ec + es);
    // e^{-i ph} = (-i)^q e^{-i rd}: the quadrant is applied to the operand in
    // fac_mul (exactly), the factor keeps the reduced angle (|cos| >= |sin|)
    Fac<float> f;
    f.q = q & 3;
    const float he = -0.5f * e;
    // with g = 1 the products below are exact and reduce to the undamped pair
    const float hx = g * cr, hy = -(g * sr);
    f.hi = mk<float>(hx, hy);
    f.lo = mk<float>(fmaf(g, cr, -hx) + he * hx, -fmaf(g, sr, hy) + he * hy);
    return f;
}
// FP64 angle (kinetic and field phases, whose angles reach many turns)
template <> __device__ __forceinline__ Fac<float> phase_factor<float>(double arg, double ph)
{
    const double j = rint(ph * 0.63661977236758134307553505349006);          // 2/pi
    double rd = fma(-j, 1.5707963267948965579989817342721, ph);               // pi/2 hi
    rd = fma(-j, 6.1232339957367660359e-17, rd);                             // pi/2 lo
    // the damping factor is computed for every point and selected (no branch)
    const float gx = exp_f32(arg > QND_MAX_EXP_ARG_F32 ? QND_MAX_EXP_ARG_F32 : arg);
    return fac_octant_f32((float)rd, (int)j, (arg == 0.0) ? 1.0f : gx);
}
// FP32 angle, for the complex64 potential phases: the angle is formed from
// the FP32 potential in FP32 (a relative error of a few 1e-8 of an angle of
// order one, the size of the potential's own rounding) and reduced by pi/2
// in three FP32 parts (Cody-Waite; exact products inside the FMAs).
__device__ __forceinline__ Fac<float> phase_factor_f32(float arg, float ph)
{
    const float j = rintf(ph * 0.636619772f);                // 2/pi
    float rd = fmaf(-j, 1.57079637050628662109375f, ph);      // pi/2 = hi + mid + lo
    rd = fmaf(-j, -4.371138828673793e-8f, rd);
    rd = fmaf(-j, -1.7151245100058819e-15f, rd);
    const float gx = exp2f((arg > (float)QND_MAX_EXP_ARG_F32 ? (float)QND_MAX_EXP_ARG_F32 : arg) * 1.44269504f);
    return fac_octant_f32(rd, (int)j, (arg == 0.0f) ? 1.0f : gx);
}
// Potential factor e^{damp} (cos ph - i sin ph).  complex128: FP64
// throughout.  complex64: the damping in FP32; the angle either formed and
// reduced mod pi/2 in FP64 (QND_POT_ANGLE_FP64, the default: c h (V + x.E)
// with V in FP32, the product, sum and reduction in FP64, sin and cos from
// the FP32 polynomials of the octant) or formed and reduced in FP32.
#ifndef QND_POT_ANGLE_FP64
#define QND_POT_ANGLE_FP64 1
#endif
// QND_POT_F64: the complex64 passes evaluate V and dV/du in FP64 (else FP32)
#ifndef QND_POT_F64
#define QND_POT_F64 1
#endif
__device__ __forceinline__ Fac<double> pot_factor(double damp, double ph) { return phase_factor<double>(damp, ph); }
__device__ __forceinline__ Fac<float> pot_factor(float damp, float ph) { return phase_factor_f32(damp, ph); }
__device__ __forceinline__ Fac<float> pot_factor(float damp, double ph)
{
    const double j = rint(ph * 0.63661977236758134307553505349006);          // 2/pi
    double rd = fma(-j, 1.5707963267948965579989817342721, ph);               // pi/2 hi
    rd = fma(-j, 6.1232339957367660359e-17, rd);                             // pi/2 lo
    const float gx = exp2f((damp > (float)QND_MAX_EXP_ARG_F32 ? (float)QND_MAX_EXP_ARG_F32 : damp) * 1.44269504f);
    return fac_octant_f32((float)rd, (int)j, (damp == 0.0f) ? 1.0f : gx);
}
// A tabulated unit factor (cos, -sin) as a Fac: complex64 tables hold the
// FP64 value as an FP32 hi + lo pair (x, y, z, w) = (hi.re, hi.im, lo.re,
// lo.im); conj gives the adjoint's conjugated factor.
template <typename R> __device__ __forceinline__ Fac<R> tab_fac(const typename TwT<R>::T &v, bool conj);
template <> __device__ __forceinline__ Fac<double> tab_fac<double>(const double2 &v, bool conj)
{
    Fac<double> f;
    f.v = mk<double>(v.x, conj ? -v.y : v.y);
    return f;
}
template <> __device__ __forceinline__ Fac<float> tab_fac<float>(const float4 &v, bool conj)
{
    // (x, y) the factor, conjugated if asked; with |y| > |x| it is stored as
    // (-i)^1 times i (x + i y) = (-y, x), whose first component is the larger
    const float x = v.x, y = conj ? -v.y : v.y, xl = v.z, yl = conj ? -v.w : v.w;
    const bool rot = fabsf(y) > fabsf(x);
    Fac<float> f;
    f.q = rot ? 1 : 0;
    f.hi = mk<float>(rot ? -y : x, rot ? x : y);
    f.lo = mk<float>(rot ? -yl : xl, rot ? xl : yl);
    return f;
}
// Factors of the complex64 storage, FP64 arithmetic form: formed as for
// complex64 (FP64 angle reduction, FP32 polynomials, modulus-corrected hi +
// lo; or a complex64 table entry) and applied as one FP64 product.
template <typename CR> struct FacCvt;
template <> struct FacCvt<float> {
    static __device__ __forceinline__ Fac<float> f(const Fac<float> &a) { return a; }
};
template <> struct FacCvt<double> {
    static __device__ __forceinline__ Fac<double> f(const Fac<float> &a)
    {
        // (-i)^q (hi + lo): q = 0 (u, v), 1 (v, -u), 2 (-u, -v), 3 (-v, u)
        const double hx = (double)a.hi.x + (double)a.lo.x, hy = (double)a.hi.y + (double)a.lo.y;
        const bool odd = (a.q & 1) != 0;
        const double u = odd ? hy : hx, v = odd ? hx : hy;
        Fac<double> r;
        r.v = mk<double>((a.q == 2 || a.q == 3) ? -u : u, (a.q == 1 || a.q == 2) ? -v : v);
        return r;
    }
};
// the phase factor e^{arg}(cos ph - i sin ph) of a pass whose state is stored
// as C<R> and held as C<CR>
template <typename R, typename CR> __device__ __forceinline__ Fac<CR> phase_fac(double arg, double ph)
{
    if constexpr (std::is_same<R, CR>::value) return phase_factor<R>(arg, ph);
    else return FacCvt<CR>::f(phase_factor<R>(arg, ph));
}
// a table entry TE as a factor of arithmetic type CR
template <typename CR, typename TE> __device__ __forceinline__ Fac<CR> tab_fac_x(const TE &v, bool conj)
{
    if constexpr (std::is_same<TE, typename TwT<CR>::T>::value) {
        return tab_fac<CR>(v, conj);
    } else {
        const double x = (double)v.x + (double)v.z, y = (double)v.y + (double)v.w;
        Fac<double> f;
        f.v = mk<double>(x, conj ? -y : y);
        return f;
    }
}
// a stored value widened to the registers' type, and narrowed back for the store
template <typename CR, typename V> __device__ __forceinline__ C<CR> cvtc(const V &v)
{
    return mk<CR>((CR)v.x, (CR)v.y);
}

// Store the unit factor (c, s) into a table entry (FP32 hi + lo, or FP64).
__device__ __forceinline__ void tab_store(float4 *t, double c, double s)
{
    const float ch = (float)c, sh = (float)s;
    *t = make_float4(ch, sh, (float)(c - (double)ch), (float)(s - (double)sh));
}
__device__ __forceinline__ void tab_store(double2 *t, double c, double s) { *t = make_double2(c, s); }

template <typename R> __device__ __forceinline__ R real_exp(double arg);
template <> __device__ __forceinline__ double real_exp<double>(double arg)
{
    return exp(arg > QND_MAX_EXP_ARG ? QND_MAX_EXP_ARG : arg);
}
template <> __device__ __forceinline__ float real_exp<float>(double arg)
{
    return exp_f32(arg > QND_MAX_EXP_ARG_F32 ? QND_MAX_EXP_ARG_F32 : arg);
}

// ---------------------------------------------------------------------
// wavefront / workgroup FP64 reductions (fixed butterfly order)
// ---------------------------------------------------------------------
// Butterfly sum over the active lanes of a wavefront: all 64 when the block
// has >= 64 threads, else the block's power-of-two size.
__device__ __forceinline__ double wave_sum(double v)
{
    const int width = blockDim.x < 64 ? (int)blockDim.x : 64;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        if (o < width) v += __shfl_xor(v, o, 64);
    return v;
}
// Reduce v over the block; thread 0 gets the result.  red = LDS scratch of
// at least 16 doubles.  All threads must call.
__device__ __forceinline__ double block_sum(double v, double *red)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    lds_barrier();
    if (lane == 0) red[w] = v;
    lds_barrier();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < nw; ++i) s += red[i];
    return s;
}
// Reduce K values over the block in one exchange and write the K sums to
// out[0..K) (thread 0).  red = LDS scratch of at least 16*K doubles.  All
// threads must call.
template <int K> __device__ __forceinline__ void block_sum_vec(double *v, double *red, double *out)
{
#pragma unroll
    for (int q = 0; q < K; ++q) v[q] = wave_sum(v[q]);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    lds_barrier();
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < K; ++q) red[q * 16 + w] = v[q];
    }
    lds_barrier();
    for (int q = threadIdx.x; q < K; q += blockDim.x) {  // blocks may be smaller than K
        double s = 0.0;
        for (int i = 0; i < nw; ++i) s += red[q * 16 + i];
        out[q] = s;
    }
}
// Sum column q of a [nwg][npart] partials array with the threads of one
// wavefront (lane-strided, then butterfly): deterministic for a given nwg.
__device__ __forceinline__ double wave_sum_column(const double *part, int nwg, int npart, int q)
{
    const int lane = threadIdx.x & 63;
    const int width = blockDim.x < 64 ? (int)blockDim.x : 64;
    double s = 0.0;
    for (int w = lane; w < nwg; w += width) s += part[(int64_t)w * npart + q];
    return wave_sum(s);
}

// complex128 value rounded to complex64 and back (PF_ROUND32); identity in
// complex64
template <typename R> __device__ __forceinline__ C<R> round_to_c64(C<R> v)
{
    if constexpr (std::is_same<R, double>::value) return mk<double>((double)(float)v.x, (double)(float)v.y);
    else return v;
}

// ---------------------------------------------------------------------
// pass arguments
// ---------------------------------------------------------------------
// PK_KINS is the kinetic pass of a merged step boundary: after the factor and
// the inverse transform it [stores the state to `save` (PF_SAVE)], [applies
// its axis's mask (PF_MASK)], and runs a forward transform, its axis's factor
// from `atab` and the inverse.  A kernel of its own, so that the second
// transform's registers stay out of PK_KIN.
// PK_KINR is its adjoint in reverse order (the costate sweep): forward
// transform, its axis's conjugated factor from `atab`, inverse, [its axis's
// mask], plus the boundary's source term from `save` (axis-0 Fourier
// representation), then the ordinary kinetic pass.
// PK_POTL is the potential pass of a length-gauge stage (flags PF_V1 |
// PF_SECOND only): PK_POT with its flags and potential mode known at compile
// time, so that none of the other operators' code is in its kernel.
// PK_KINP is the kinetic pass of the phase mode (length gauge) that finalizes
// no record: PK_KIN with its mode known at compile time.  PK_POTI and PK_KINI
// are the same for imaginary time: the potential pass between two stages
// (PF_V1 | PF_V2 | PF_SECOND) and the kinetic pass.
// PK_POTX is the first potential pass after a merged boundary
// (PF_V1 | PF_SECOND | PF_PREKX, [PF_MASK]), length gauge; PK_POTO the
// batched record pass (PF_OBS | PF_NOSTORE, [PF_MASK]).
// PK_POTS and PK_POTG are the adjoint's length-gauge stage passes: the
// recompute's (PF_V1 | PF_SAVE | PF_SECOND) and the costate sweep's
// (PF_V1 | PF_CONJ | PF_GRAD | PF_SECOND).
// PK_POTV and PK_KINV are the velocity gauge's stage potential pass
// (PF_V1 | PF_V2 | PF_SECOND) and kinetic pass.
// PK_P1DI is the 1-D imaginary-time stage pass (PF_V1 | PF_V2 | PF_SECOND |
// PF_KPOST, both modes imaginary).
// PK_POTJ is the imaginary-time potential pass with its flags from the launch,
// among PF_V1, PF_V2, PF_SECOND and the Gram-Schmidt update and dots (the
// first and last potential pass of a step, dim >= 2); PK_POTB the observable
// boundary pass (PF_OBS | PF_SECOND, [PF_MASK]: owed mask and records, no
// potential factor).
enum PassKind { PK_PLAIN = 0, PK_KIN = 1, PK_POT = 2, PK_POT1D = 3, PK_KINS = 4, PK_KINR = 5, PK_POTL = 6,
                PK_KINP = 7, PK_POTI = 8, PK_KINI = 9, PK_POTX = 10, PK_POTO = 11, PK_POTS = 12, PK_POTG = 13,
                PK_POTV = 14, PK_KINV = 15, PK_P1DI = 16, PK_POTJ = 17, PK_POTB = 18 };
enum PassFlags : unsigned {
    PF_GS_UPDATE = 1u << 0,  // prologue: v = (v - sum_j c_j prev_j) * invnorm
    PF_KPRE = 1u << 1,       // (dim = 1) kinetic factor before the first transform
    PF_V1 = 1u << 2,         // potential factor #1 (after the first transform)
    PF_MASK = 1u << 3,       // cos^(1/8) mask
    PF_OBS = 1u << 4,        // r-space observable partial sums
    PF_V2 = 1u << 5,         // potential factor #2
    PF_SECOND = 1u << 6,     // second (reverse) transform
    PF_KPOST = 1u << 7,      // (dim = 1) kinetic factor after the second transform
    PF_GS_DOTS = 1u << 8,    // Gram-Schmidt partial dots after the second transform
    PF_FINALIZE = 1u << 9,   // extra workgroups turn the previous pass's observable partials into records
    // adjoint (costate) passes
    PF_SAVE = 1u << 10,      // store the position-space state after the r-space operators into `save`
    PF_CONJ = 1u << 11,      // conjugate phase factors (damping unchanged): the adjoint of the pass
    PF_GRAD = 1u << 12,      // partials of x_j Im(conj(u) phi) and dV/da Im(conj(u) phi), phi from `save`
    PF_SRC = 1u << 13,       // u += sum_j src_j (w_j - E_j) y + src_norm y (y from `save`), w = -grad V
    PF_KTAB = 1u << 14,      // kinetic factor read from `ktab` (tabulated e^{-i c K}, four values per thread)
    PF_KAX = 1u << 15,       // kinetic factor from per-axis tables in `ktab` (phase: e^{-i c K_j}; imaginary
                             // time: e^{-c K_j} in the first component), Sigma n_j entries at the per-axis
                             // offsets: this axis's entry per element times the folded axes' entries of the line
    // merged step boundary (2-D length gauge): K(a_last) | record | Mask | K(a_0);
    // its kinetic pass is the kernel kind PK_KINS
    PF_PREKX = 1u << 16,     // axis-0 pass: after the first transform, [axis-0 mask], forward, axis-0
                             // factor from `atab`, inverse, then the r-space operators
    PF_NOSTORE = 1u << 17,   // no store (observable-only pass)
    PF_SRCM = 1u << 18,      // with PF_SRC: `save` holds the record's state in the axis-0 Fourier
                             // representation (merged boundary); y = [Mask] F0^-1 save / n0
    PF_ZERO = 1u << 19,      // with PF_SRC: the costate is zero (the pass forms the source term alone)
    PF_POSTKX = 1u << 20,    // axis-0 pass: after the second transform, the conjugated axis-0 factor
                             // from `atab`, inverse, [axis-0 mask], forward (adjoint of PF_PREKX)
    PF_ROUND32 = 1u << 21,   // complex128 passes: round every stored state value to complex64 (a
                             // precision experiment: FP64 arithmetic with complex64 storage)
};
enum KMode { KM_PHASE = 0, KM_VELOCITY = 1, KM_IMAG = 2 };
enum VMode { VM_LENGTH = 0, VM_VELOCITY = 1, VM_IMAG = 2 };

#define QND_NOBS (1 + 2 * QND_MAXD)

// observable record assembled from r-space partial sums [n0, d_j, g_j]
struct ObsSpec {
    int ncodes;
    int codes[8];
    double E[QND_MAXD];
};

// Batched instances: nb consecutive slots propagate in one launch, each under
// its own field.  Field rows are FP64 [dim] vectors; instance b's row is at
// base + b * stride.  The observable record of instance b is finalized from
// the partials of its own workgroups (they never straddle instances).
struct Batch {
    int nb;                  // instances (1: the scalar fields of PassArgs apply)
    int64_t lpi;             // lines per instance of this pass
    const double *fld;       // this stage's E (length) or A (velocity) of instance 0
    int64_t fld_stride;
    const double *finE;      // E(t_n) of instance 0 for the record being finalized
    int64_t finE_stride;
    int64_t rec_stride;      // doubles between the records of two instances
    const double *srcr;      // PF_SRC: per-instance source weights [dim] (nullptr: PassArgs::src)
    int64_t srcr_stride;
    // Step counter on the device (nullptr: none): the field rows, E(t_n) and
    // the record then advance with *step, so that one launch sequence serves
    // every step (a replayed graph): row = fld + *step * fld_step, E(t_n) =
    // finE + *step * finE_step, record = rec + (*step - step_base) * rec_step.
    const int64_t *step;
    int64_t fld_step, finE_step, rec_step, step_base;
};

template <typename R> struct PassArgs {
    C<R> *data;
    const typename TwT<R>::T *tw;
    int axis, n, log2n, T, log2T, TW, log2TW, ldsStride;
    int wave_lines;          // allow wavefront-local exchanges for 64-thread contiguous lines
    int64_t S, nlines;
    unsigned flags;
    double scale1, scale2;   // after the first / second transform
    int kmode;
    unsigned kin_axes;       // kinetic pass: the axes whose kinetic terms it applies (always its own)
    double ck_pre, ck_mid, ck_post;
    double Ak[QND_MAXD];
    int vmode;
    double cv1, cv2;
    double Ev[QND_MAXD];
    // kinetic factors e^{-i c K} of this pass's coefficient at every element of
    // one instance (row-major), when tabulated (PF_KTAB); or the per-axis
    // factors of this pass's coefficient (PF_KAX)
    const typename TwT<R>::T *ktab;
    // per-axis kinetic factors e^{-i c K_j} of one coefficient (the step's
    // first, a_0), Sigma n_j entries at the per-axis offsets (PK_KINS, PF_PREKX)
    const typename TwT<R>::T *atab;
    // V(r) in FP64 at every element of one instance (row-major), when
    // tabulated (four values per thread); nullptr: evaluated in the pass
    const double *vtab;
    // adjoint passes: forward states saved in position space, source weights
    C<R> *save;
    double src[QND_MAXD];
    double src_norm;
    // Gram-Schmidt
    const C<R> *prev;
    int64_t slot_stride;
    int nprev;
    const double *gs_coef;   // [2*nprev] c_j, invnorm, flag
    double *partials;        // per-workgroup outputs of this pass
    int npart;
    // observable finalize (of the previous pass's partials)
    const double *fin_part;
    int fin_nwg;
    double *fin_rec;
    ObsSpec fin_spec;
    Batch B;
    Phys P;
};


// Multi-index of the line L (all axes but `axis`).  L = o * S + in, with
// `in` the index over the axes after `axis` and `o` over those before it (a
// batch index above axis 0 drops out of the modulo).
__device__ __forceinline__ void line_index(const Phys &P, int axis, int64_t S, int log2S, int64_t L, int *idx)
{
    if (P.pow2) {
        int64_t o = L >> log2S, in = L & (S - 1);
        int sh = 0;
#pragma unroll
        for (int j = QND_MAXD - 1; j >= 0; --j) {
            idx[j] = 0;
            if (j < P.dim && j > axis) {
                idx[j] = (int)((in >> sh) & (P.n[j] - 1));
                sh += P.lg[j];
            }
        }
        sh = 0;
#pragma unroll
        for (int j = QND_MAXD - 1; j >= 0; --j) {
            if (j < axis) {
                idx[j] = (int)((o >> sh) & (P.n[j] - 1));
                sh += P.lg[j];
            }
        }
        return;
    }
    int64_t o = L / S, in = L - o * S;
#pragma unroll
    for (int j = QND_MAXD - 1; j >= 0; --j) {
        idx[j] = 0;
        if (j < P.dim && j > axis) {
            idx[j] = (int)(in % P.n[j]);
            in /= P.n[j];
        }
    }
#pragma unroll
    for (int j = QND_MAXD - 1; j >= 0; --j) {
        if (j < axis) {
            idx[j] = (int)(o % P.n[j]);
            o /= P.n[j];
        }
    }
}

// Column sums of a [nwg][npart] partials array, columns q0 .. q0+K-1, with
// every thread of the block: thread-strided partial sums, then the fixed
// tree of block_reduce_lds.  Deterministic for a given block size.  All
// threads must call; every thread gets the sums in tot[].
__device__ __forceinline__ void block_reduce_lds(double *v, int K, double *red);
template <int K>
__device__ __forceinline__ void block_column_sums(const double *part, int nwg, int npart, int q0, double *tot,
                                                  double *red)
{
    double acc[K];
#pragma unroll
    for (int q = 0; q < K; ++q) acc[q] = 0.0;
    for (int w = threadIdx.x; w < nwg; w += blockDim.x) {
#pragma unroll
        for (int q = 0; q < K; ++q)
            if (q0 + q < npart) acc[q] += part[(int64_t)w * npart + q0 + q];
    }
    block_reduce_lds(acc, K, red);
#pragma unroll
    for (int q = 0; q < K; ++q) tot[q] = acc[q];
}

// Sums v[0..K) over the block (wave butterflies, then the waves in order);
// every thread receives the result in v.  red: LDS of >= 16*K doubles.
__device__ __forceinline__ void block_reduce_lds(double *v, int K, double *red)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    for (int q = 0; q < K; ++q) v[q] = wave_sum(v[q]);
    lds_barrier();
    if (lane == 0)
        for (int q = 0; q < K; ++q) red[q * 16 + w] = v[q];
    lds_barrier();
    for (int q = 0; q < K; ++q) {
        double s = 0.0;
        for (int i = 0; i < nw; ++i) s += red[q * 16 + i];
        v[q] = s;
    }
    lds_barrier();
}

// Turn per-workgroup [n0, d_j, g_j] sums into the observable record.  Every
// thread of the calling block must call (block-wide reduction).
__device__ __forceinline__ void obs_finalize(const double *part, int nwg, int dim, double dNr,
                                             const ObsSpec &spec, const double *E, double *rec, double *red)
{
    double tot[QND_NOBS];
    block_column_sums<QND_NOBS>(part, nwg, QND_NOBS, 0, tot, red);
#pragma unroll
    for (int q = 0; q < QND_NOBS; ++q) tot[q] *= dNr;
    if (threadIdx.x == 0) {
        int off = 0;
        for (int k = 0; k < spec.ncodes; ++k) {
            const int c = spec.codes[k];
            if (c == 5) rec[off++] = tot[0];                        // norm
            else if (c == 3) {                                      // dipole_r
#pragma unroll
                for (int j = 0; j < QND_MAXD; ++j)
                    if (j < dim) rec[off++] = tot[1 + j];
            } else if (c == 6) {                                    // dipole_a
#pragma unroll
                for (int j = 0; j < QND_MAXD; ++j)
                    if (j < dim) rec[off++] = -tot[1 + QND_MAXD + j] - E[j] * tot[0];
            }
        }
    }
}
// Instance-b view of a finalize request: partials, E(t_n) and record.
__device__ __forceinline__ void obs_finalize_instance(const double *part, int nwg, int dim, double dNr,
                                                      const ObsSpec &spec, const Batch &B, int b, double *rec,
                                                      double *red)
{
    const int64_t sd = B.step ? *B.step : 0;
    const double *E = B.finE ? B.finE + sd * B.finE_step + b * B.finE_stride : spec.E;
    obs_finalize(part + (int64_t)b * nwg * QND_NOBS, nwg, dim, dNr, spec, E,
                 rec + (sd - B.step_base) * B.rec_step + b * B.rec_stride, red);
}


// ---------------------------------------------------------------------
// the line-pass kernel
// ---------------------------------------------------------------------
#define QND_PASS_THREADS 512   // workgroup cap: leaves 256 VGPRs per lane
// Gram-Schmidt in the passes: previous states whose values are in flight at
// once (one memory latency per group), and states per LDS chunk of the dots
#define QND_GS_G(R) (sizeof(R) == 4 ? 2 : 1)
#define QND_GS_CHUNK 8
// Touch every 64-byte line of the kernel's argument block at entry, so that
// their cache misses overlap.  The compiler reads arguments lazily, near
// their first use and again later; each first read of a line would
// otherwise be a separate round trip to memory (600-1900 cycles measured
// on the line passes) on the critical path of every wavefront.
template <int BYTES> __device__ __forceinline__ void kernarg_touch()
{
    const __attribute__((address_space(4))) unsigned *k =
        (const __attribute__((address_space(4))) unsigned *)__builtin_amdgcn_kernarg_segment_ptr();
    constexpr int NL = (BYTES + 63) / 64;
    static_assert(NL <= 13, "argument block larger than the touched lines");
    unsigned v[13];
#pragma unroll
    for (int i = 0; i < 13; ++i) v[i] = k[16 * (i < NL ? i : 0)];
    // one consumer of all values: every load is issued before the one wait
    asm volatile("" ::"s"(v[0]), "s"(v[1]), "s"(v[2]), "s"(v[3]), "s"(v[4]), "s"(v[5]), "s"(v[6]), "s"(v[7]),
                 "s"(v[8]), "s"(v[9]), "s"(v[10]), "s"(v[11]), "s"(v[12]));
}

// The pass kernel's leading scalar arguments are what the addresses of its
// first loads need: the state and twiddle pointers, the element stride S of
// a line (1: contiguous lines), the packed shape `geo` (below), the flags
// and the number of workgroups (gridDim.x would be read from the argument
// block).
// Built with kernel-argument preloading, they arrive in SGPRs with the
// wavefront, so the state and twiddle loads go out before the argument block
// (PassArgs) has been read.
// geo = log2n | log2T << 4 | log2TW << 8 | pow2 << 12
__host__ __device__ inline unsigned pass_geo(int log2n, int log2T, int log2TW, int pow2)
{
    return (unsigned)log2n | ((unsigned)log2T << 4) | ((unsigned)log2TW << 8) | ((unsigned)(pow2 != 0) << 12);
}

// The body of a line pass: workgroup `bid` of the pass's nblk line tiles
// (and, for a finalizing kinetic pass, of the finalize workgroups after
// them).  TOUCH: touch the argument block's lines at entry (the kernel's
// PassArgs is its only argument block).
template <typename R, int KIND, int DIR1, int VPT, bool TOUCH, bool X64 = false>
__device__ __forceinline__ void pass_body(C<R> *data, const typename TwT<R>::T *tw, int64_t S, unsigned geo,
                                          unsigned fl_arg, int nblk, const PassArgs<R> &a, int bid)
{
    // X64: the state is stored as C<R> (complex64) and held in FP64 registers
    // (loads widen, the store rounds once); every pass operator runs in CR
    using CR = typename std::conditional<X64, double, R>::type;
    const unsigned fl = KIND == PK_POTL   ? (unsigned)(PF_V1 | PF_SECOND)
                        : (KIND == PK_POTI || KIND == PK_POTV) ? (unsigned)(PF_V1 | PF_V2 | PF_SECOND)
                        : KIND == PK_P1DI ? (unsigned)(PF_V1 | PF_V2 | PF_SECOND | PF_KPOST)
                        : KIND == PK_POTX ? (unsigned)(PF_V1 | PF_SECOND | PF_PREKX) | (fl_arg & PF_MASK)
                        : KIND == PK_POTO ? (unsigned)(PF_OBS | PF_NOSTORE) | (fl_arg & PF_MASK)
                        : KIND == PK_POTS ? (unsigned)(PF_V1 | PF_SAVE | PF_SECOND)
                        : KIND == PK_POTG ? (unsigned)(PF_V1 | PF_CONJ | PF_GRAD | PF_SECOND)
                        : KIND == PK_POTJ ? fl_arg & (unsigned)(PF_V1 | PF_V2 | PF_SECOND | PF_GS_UPDATE | PF_GS_DOTS)
                        : KIND == PK_POTB ? (unsigned)(PF_OBS | PF_SECOND) | (fl_arg & PF_MASK)
                                          : fl_arg;
    const int vmode = (KIND == PK_POTL || KIND == PK_POTX || KIND == PK_POTS || KIND == PK_POTG) ? (int)VM_LENGTH
                      : (KIND == PK_POTI || KIND == PK_P1DI || KIND == PK_POTJ)                ? (int)VM_IMAG
                      : KIND == PK_POTV                                                        ? (int)VM_VELOCITY
                                                                                               : a.vmode;
    const int kmode = (KIND == PK_KINP || KIND == PK_KINS || KIND == PK_KINR) ? (int)KM_PHASE
                      : (KIND == PK_KINI || KIND == PK_P1DI)                  ? (int)KM_IMAG
                      : KIND == PK_KINV                                       ? (int)KM_VELOCITY
                                                                              : a.kmode;
    // complex128 only: the stored state rounded to complex64 (experiment switch)
    const bool r32 = std::is_same<R, double>::value && (fl_arg & PF_ROUND32) != 0;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[16 * QND_NOBS];
    C<CR> *lds = reinterpret_cast<C<CR> *>(smem_raw);  // [2 buffers][TW lines][ldsStride]
    constexpr bool POTK = (KIND == PK_POT || KIND == PK_POT1D || KIND == PK_POTL || KIND == PK_POTI ||
                           KIND == PK_POTX || KIND == PK_POTO || KIND == PK_POTS || KIND == PK_POTG ||
                           KIND == PK_POTV || KIND == PK_P1DI || KIND == PK_POTJ || KIND == PK_POTB);
    constexpr bool KINK = (KIND == PK_KIN || KIND == PK_KINS || KIND == PK_KINR || KIND == PK_KINP ||
                           KIND == PK_KINI || KIND == PK_KINV);

    // workgroups after the nblk line tiles finalize the previous pass's
    // observable record (one per instance) and do no line work, so the
    // record does not delay a tile
    if (KIND == PK_KIN && bid >= nblk) {
        const int b = bid - nblk;
        if ((fl & PF_FINALIZE) && b < a.B.nb)
            obs_finalize_instance(a.fin_part, a.fin_nwg, a.P.dim, a.P.dNr, a.fin_spec, a.B, b, a.fin_rec, red);
        return;
    }
    const int tid = threadIdx.x;
    const int log2n = geo & 15, log2T = (geo >> 4) & 15, log2TW = (geo >> 8) & 15;
    const bool pow2 = (geo >> 12) & 1;
    const int n = (VPT == 12 ? 3 : 1) << log2n, T = 1 << log2T, TW = 1 << log2TW;
    const Phys &P = a.P;

    // ---- tile geometry: TW lines (TW | S or S == 1).  Contiguous lines put
    // t on the fastest lanes, strided lines the line index l, so that every
    // wavefront access covers consecutive addresses. ----
    const bool contig = (S == 1);
    int log2S = 0;
    while (((int64_t)1 << log2S) < S) ++log2S;
    int l, t;
    if (contig) { l = tid >> log2T; t = tid & (T - 1); }
    else { l = tid & (TW - 1); t = tid >> log2TW; }
    // XCD-aware tile order: workgroups are dealt round-robin to the 8 XCDs,
    // so workgroup w runs on XCD w % 8.  Tile tb = (w % 8) * (grid / 8) + w / 8
    // gives each XCD a contiguous run of tiles: neighbouring strided tiles,
    // which share cache lines, then share one L2 instead of fetching the same
    // lines into several.
    const int tb = (nblk % 8 == 0) ? (bid % 8) * (nblk / 8) + bid / 8 : bid;
    const int64_t L = (int64_t)tb * TW + l;  // this thread's line
    int64_t lbase, estride;                          // element i at lbase + i * estride
    if (contig) { lbase = L * n; estride = 1; }
    else if (pow2) { lbase = (((L >> log2S) << log2n) << log2S) + (L & (S - 1)); estride = S; }
    else { const int64_t o = L / S; lbase = o * n * S + (L - o * S); estride = S; }
    Tw4<CR> tw4;  // four values per thread: all twiddles requested before the state
    if constexpr (VPT == 4) tw4_load<CR>(tw4, tw, log2n, T, t);

    // ---- load x[m] = element t + m T (with the Gram-Schmidt update prologue).
    // Twiddles are read from the table in memory (L1/L2 hits); staging it in
    // LDS measured slower. ----
    C<CR> x[VPT];
#pragma unroll
    for (int m = 0; m < VPT; ++m) x[m] = cvtc<CR>(data[lbase + (int64_t)(t + m * T) * estride]);
    // the argument block, now that the first loads are out
    if constexpr (TOUCH) kernarg_touch<sizeof(PassArgs<R>) + 40>();
    const int ldsS = a.ldsStride;
    C<CR> *bufA = lds + l * ldsS;
    C<CR> *bufB = lds + (TW + l) * ldsS;
    int wc = 0;

    if (POTK && (fl & PF_GS_UPDATE)) {
        const double gsinv = a.gs_coef[2 * a.nprev];
        double vr[VPT], vi[VPT];
#pragma unroll
        for (int m = 0; m < VPT; ++m) { vr[m] = (double)x[m].x; vi[m] = (double)x[m].y; }
        // the previous states' values are requested QND_GS_G states at a
        // time, so that one memory latency serves several states (the
        // subtraction order stays j = 0, 1, ...)
        constexpr int G = QND_GS_G(CR);
        for (int j0 = 0; j0 < a.nprev; j0 += G) {
            C<R> p[G][VPT];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (j0 + g < a.nprev) {
                    const C<R> *pj = a.prev + (int64_t)(j0 + g) * a.slot_stride;
#pragma unroll
                    for (int m = 0; m < VPT; ++m) p[g][m] = pj[lbase + (int64_t)(t + m * T) * estride];
                }
            }
#pragma unroll
            for (int g = 0; g < G; ++g) {
                if (j0 + g < a.nprev) {
                    const double cr = a.gs_coef[2 * (j0 + g)], ci = a.gs_coef[2 * (j0 + g) + 1];
#pragma unroll
                    for (int m = 0; m < VPT; ++m) {
                        vr[m] -= cr * (double)p[g][m].x - ci * (double)p[g][m].y;
                        vi[m] -= cr * (double)p[g][m].y + ci * (double)p[g][m].x;
                    }
                }
            }
        }
#pragma unroll
        for (int m = 0; m < VPT; ++m) x[m] = mk<CR>((CR)(vr[m] * gsinv), (CR)(vi[m] * gsinv));
    }

    const int ax = a.axis;
    const double cs = (fl & PF_CONJ) ? -1.0 : 1.0;  // adjoint: conjugated phases
    // a contiguous line of exactly 64 threads is one wavefront: its exchanges
    // need no workgroup barrier
    const bool wave_line = contig && T == 64 && a.wave_lines;
    if (KIND == PK_PLAIN) {
        transform<CR, DIR1, VPT>(x, bufA, bufB, wc, n, log2n, T, t, tw, tw4, wave_line);
        if (a.scale1 != 1.0) {
            const CR s1 = (CR)a.scale1;
#pragma unroll
            for (int m = 0; m < VPT; ++m) x[m] = scl<CR>(x[m], s1);
        }
    } else {
        int idx[QND_MAXD];
        line_index(P, ax, S, log2S, L, idx);
        // per-axis CAP and mask values of this thread's elements (both tables
        // always exist).  With four values per thread (one wavefront per
        // SIMD, latency-bound) they are requested before the transform so
        // that the loads overlap it; with more values per thread the passes
        // hide latency with occupancy, and each value is read where it is
        // used, which keeps 2 VPT doubles out of the registers of the
        // transforms and of the other points.
        double capv[VPT], mskv[VPT];
        double Vco = 0.0, mko = 1.0;  // CAP and mask of the other axes (same order of terms)
        if (POTK) {
            if constexpr (VPT == 4) {
#pragma unroll
                for (int m = 0; m < VPT; ++m) {
                    capv[m] = P.capt[P.off[ax] + t + m * T];
                    mskv[m] = P.maskt[P.off[ax] + t + m * T];
                }
            }
#pragma unroll
            for (int j = 0; j < QND_MAXD; ++j) {
                if (j < P.dim && j != ax) {
                    if (P.cap_on) Vco += P.capt[P.off[j] + idx[j]];
                    if (P.mask_on) mko *= P.maskt[P.off[j] + idx[j]];
                }
            }
        }
        auto cap_at = [&](int m) { return VPT == 4 ? capv[m] : P.capt[P.off[ax] + t + m * T]; };
        auto msk_at = [&](int m) { return VPT == 4 ? mskv[m] : P.maskt[P.off[ax] + t + m * T]; };
        // this instance's stage field (E in the length gauge, A in the
        // velocity gauge): kernel arguments, or a row of the batch tables
        // requested here so that the load overlaps the transform
        double fv[QND_MAXD];
        {
            const bool use_a = KINK || kmode == KM_VELOCITY;
            if (a.B.fld) {
                const int64_t sd = a.B.step ? *a.B.step : 0;
                const double *src = a.B.fld + sd * a.B.fld_step + (L / a.B.lpi) * a.B.fld_stride;
#pragma unroll
                for (int j = 0; j < QND_MAXD; ++j) fv[j] = (j < P.dim) ? src[j] : 0.0;
            } else {
#pragma unroll
                for (int j = 0; j < QND_MAXD; ++j) fv[j] = (j < P.dim) ? (use_a ? a.Ak[j] : a.Ev[j]) : 0.0;
            }
        }

        // tabulated potential of this thread's elements (four values per
        // thread), requested before the transform
        double vtv[VPT];
        if constexpr (POTK && VPT == 4) {
            if (a.vtab) {
                const int64_t ib = lbase - (L / a.B.lpi) * ((int64_t)a.B.lpi * n);  // instance-relative
#pragma unroll
                for (int m = 0; m < VPT; ++m) vtv[m] = a.vtab[ib + (int64_t)(t + m * T) * estride];
            }
        }

        // tabulated kinetic factors of this thread's elements (contiguous
        // lines of the last axis), requested before the transform
        typename TwT<R>::T ktv[VPT];
        if constexpr (KINK && VPT == 4) {
            if (fl & PF_KTAB) {
                const int64_t e0 = (L % a.B.lpi) * n + t;
#pragma unroll
                for (int m = 0; m < VPT; ++m) ktv[m] = a.ktab[e0 + m * T];
            }
        }

        // merged step boundary: per-axis kinetic factors (and the kinetic
        // pass's mask values) of this thread's elements; with four values per
        // thread requested before the transform, else read where used
        typename TwT<R>::T atv[VPT];
        double amv[VPT];
        const bool split = KIND == PK_KINS || KIND == PK_KINR || KIND == PK_POTX ||
                           (KIND == PK_POT && (fl & (PF_PREKX | PF_POSTKX)));
        if constexpr (VPT == 4) {
            if (split) {
#pragma unroll
                for (int m = 0; m < VPT; ++m) atv[m] = a.atab[P.off[ax] + t + m * T];
                if ((KIND == PK_KINS || KIND == PK_KINR) && (fl & PF_MASK) && P.mask_on) {
#pragma unroll
                    for (int m = 0; m < VPT; ++m) amv[m] = P.maskt[P.off[ax] + t + m * T];
                }
            }
        }
        auto atab_at = [&](int m) { return VPT == 4 ? atv[m] : a.atab[P.off[ax] + t + m * T]; };
        auto kmask_at = [&](int m) { return VPT == 4 ? amv[m] : P.maskt[P.off[ax] + t + m * T]; };

        if ((KIND == PK_POT1D || KIND == PK_P1DI) && (fl & PF_KPRE)) {
#pragma unroll
            for (int m = 0; m < VPT; ++m) {
                const double p = coord_p(P, ax, t + m * T);
                const double K = .5 * p * p;
                if (kmode == KM_IMAG) x[m] = scl<CR>(x[m], real_exp<CR>(-a.ck_pre * K));
                else x[m] = fac_mul<CR>(x[m], phase_fac<R, CR>(0.0, cs * a.ck_pre * K));
            }
        }

        if constexpr (KIND == PK_KINR) {
            // adjoint of the merged boundary's kinetic pass, first half: this
            // axis's part of K(a_0) conjugated, this axis's mask, then the
            // boundary's source term (axis-0 Fourier representation)
#pragma unroll
            for (int m = 0; m < VPT; ++m) x[m] = scl<CR>(x[m], (CR)a.scale2);
            transform<CR, DIR1, VPT>(x, bufA, bufB, wc, n, log2n, T, t, tw, tw4, wave_line);
#pragma unroll
            for (int m = 0; m < VPT; ++m) x[m] = fac_mul<CR>(x[m], tab_fac_x<CR>(atab_at(m), cs < 0.0));
            transform<CR, -DIR1, VPT>(x, bufA, bufB, wc, n, log2n, T, t, tw, tw4, wave_line);
            const bool mk = (fl & PF_MASK) && P.mask_on;
#pragma unroll
            for (int m = 0; m < VPT; ++m) {
                const C<CR> src = cvtc<CR>(a.save[lbase + (int64_t)(t + m * T) * estride]);
                const C<CR> v = mk ? scl<CR>(x[m], (CR)kmask_at(m)) : x[m];
                x[m] = add<CR>(v, src);
            }
        }

        // ---- first transform ----
        transform<CR, DIR1, VPT>(x, bufA, bufB, wc, n, log2n, T, t, tw, tw4, wave_line);

        // adjoint source of a merged boundary: the record's state y from its
        // saved axis-0 Fourier representation (inverse transform, scale, mask)
        C<CR> ysv[VPT];
        if (KIND == PK_POT && (fl & PF_SRCM)) {
#pragma unroll
            for (int m = 0; m < VPT; ++m) ysv[m] = cvtc<CR>(a.save[lbase + (int64_t)(t + m * T) * estride]);
            transform<CR, DIR1, VPT>(ysv, bufA, bufB, wc, n, log2n, T, t, tw, tw4, wave_line);
            const bool mky = (fl & PF_MASK) && P.mask_on;
#pragma unroll
            for (int m = 0; m < VPT; ++m) {
                const double f = mky ? msk_at(m) * mko * a.scale1 : a.scale1;
                ysv[m] = mk<CR>((CR)((double)ysv[m].x * f), (CR)((double)ysv[m].y * f));
            }
        }
        if (KIND == PK_POT && (fl & PF_ZERO)) {
#pragma unroll
            for (int m = 0; m < VPT; ++m) x[m] = mk<CR>((CR)0, (CR)0);
        }

        if ((KIND == PK_POT || KIND == PK_POTX) && (fl & PF_PREKX)) {
            // the step's first kinetic factor along axis 0, after the owed
            // axis-0 mask (the kinetic pass before applied the other axes'
            // parts; the factors of different axes commute with each other
            // and with the other axes' masks)
            const bool mk0 = (fl & PF_MASK) && P.mask_on;
#pragma unroll
            for (int m = 0; m < VPT; ++m) {
                const CR f = mk0 ? (CR)(msk_at(m) * a.scale1) : (CR)a.scale1;
                x[m] = scl<CR>(x[m], f);
            }
            transform<CR, -DIR1, VPT>(x, bufA, bufB, wc, n, log2n, T, t, tw, tw4, wave_line);
#pragma unroll
            for (int m = 0; m < VPT; ++m) x[m] = fac_mul<CR>(x[m], tab_fac_x<CR>(atab_at(m), cs < 0.0));
            transform<CR, DIR1, VPT>(x, bufA, bufB, wc, n, log2n, T, t, tw, tw4, wave_line);
        }

        if (KINK) {
            // all axes in k-space: the kinetic factor of this pass's axes
            // (reference order of the sums).  exp(-i c sum_j K_j) is a product
            // of commuting per-axis factors, so in 3-D and above each pass of
            // the kinetic block applies its own axes' part.
            double Ko = 0.0, Wk = 0.0;
#pragma unroll
            for (int j = 0; j < QND_MAXD; ++j) {
                if (j < P.dim && j != ax && ((a.kin_axes >> j) & 1u)) {
                    const double pj = coord_p(P, j, idx[j]);
                    Ko += .5 * pj * pj;
                    Wk += (pj + .5 * fv[j]) * fv[j];
                }
            }
            const double Aax = fv[ax];
            // one loop per mode (mode tests outside the loop over the points)
            if (KINK && VPT == 4 && (fl & PF_KTAB)) {
#pragma unroll
                for (int m = 0; m < VPT; ++m) x[m] = fac_mul<CR>(x[m], tab_fac_x<CR>(ktv[m], cs < 0.0));
            } else if ((fl & PF_KAX) && kmode == KM_IMAG) {
                // e^{-c K} = prod_j e^{-c K_j}: this axis's factor per element,
                // times the folded axes' factors of the line
                const typename TwT<R>::T *kx = a.ktab;
                CR fo = (CR)1;
#pragma unroll
                for (int j = 0; j < QND_MAXD; ++j)
                    if (j < P.dim && j != ax && ((a.kin_axes >> j) & 1u)) fo *= (CR)kx[P.off[j] + idx[j]].x;
#pragma unroll
                for (int m = 0; m < VPT; ++m) x[m] = scl<CR>(x[m], (CR)kx[P.off[ax] + t + m * T].x * fo);
            } else if ((fl & PF_KAX) && kmode == KM_PHASE) {
                // e^{-i c K} = prod_j e^{-i c K_j}, each factor a unit table
                // entry applied by the compensated product (one rounding of
                // the state per factor)
                const typename TwT<R>::T *kx = a.ktab;
#pragma unroll
                for (int m = 0; m < VPT; ++m) x[m] = fac_mul<CR>(x[m], tab_fac_x<CR>(kx[P.off[ax] + t + m * T], cs < 0.0));
#pragma unroll
                for (int j = 0; j < QND_MAXD; ++j) {
                    if (j < P.dim && j != ax && ((a.kin_axes >> j) & 1u)) {
                        const Fac<CR> fo = tab_fac_x<CR>(kx[P.off[j] + idx[j]], cs < 0.0);
#pragma unroll
                        for (int m = 0; m < VPT; ++m) x[m] = fac_mul<CR>(x[m], fo);
                    }
                }
            } else if (kmode == KM_IMAG) {
#pragma unroll
                for (int m = 0; m < VPT; ++m) {
                    const double p = coord_p(P, ax, t + m * T);
                    const double K = Ko + .5 * p * p;
                    x[m] = scl<CR>(x[m], real_exp<CR>(-a.ck_mid * K));
                }
            } else if (kmode == KM_VELOCITY) {
#pragma unroll
                for (int m = 0; m < VPT; ++m) {
                    const double p = coord_p(P, ax, t + m * T);
                    const double K = Ko + .5 * p * p;
                    x[m] = fac_mul<CR>(x[m], phase_fac<R, CR>(0.0, cs * a.ck_mid * (K + (Wk + (p + .5 * Aax) * Aax))));
                }
            } else {
#pragma unroll
                for (int m = 0; m < VPT; ++m) {
                    const double p = coord_p(P, ax, t + m * T);
                    const double K = Ko + .5 * p * p;
                    x[m] = fac_mul<CR>(x[m], phase_fac<R, CR>(0.0, cs * a.ck_mid * K));
                }
            }
        } else {
            // all axes in r-space: potential, CAP, field, mask, observables
            double rro = 0.0, Wv = 0.0;
            double xo[QND_MAXD] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int j = 0; j < QND_MAXD; ++j) {
                if (j < P.dim && j != ax) {
                    const double xj = coord_x(P, j, idx[j]);
                    xo[j] = xj;
                    rro += xj * xj;
                    Wv += xj * fv[j];
                }
            }
            double obs[QND_NOBS];  // FP64 partials of this thread (adjoint terms, observables)
            // Per-point arithmetic runs in the precision of the state: the
            // potential, its phase angle and this thread's observable sums are
            // FP64 for complex128 and FP32 for complex64 (a thread sums its VPT
            // points in obr, which then enter the FP64 workgroup reduction).
            constexpr bool F64 = std::is_same<CR, double>::value;
            // precision of the potential: the state's, or FP64 in complex64
            // too (QND_POT_F64)
            using PR = typename std::conditional<F64 || QND_POT_F64, double, R>::type;
            R obr[QND_NOBS];
#pragma unroll
            for (int q = 0; q < QND_NOBS; ++q) {
                obs[q] = 0.0;
                obr[q] = (CR)0;
            }
            auto add_obs = [&](int q, CR val) {
                if constexpr (F64) obs[q] += val;
                else obr[q] += val;
            };
            const CR cv1 = (CR)a.cv1, cv2 = (CR)a.cv2, csr = (CR)cs, fva = (CR)fv[ax], Wvr = (CR)Wv;
            // potential factor of damping `damp` and angle c * arg: the angle
            // in FP64 (argd, cd) or, for the FP32-angle build of complex64, in
            // FP32 (argr, cr)
            auto pf = [&](CR damp, CR cr, double cd, CR argr, double argd) -> Fac<CR> {
                if constexpr (X64) return FacCvt<CR>::f(pot_factor((float)damp, cs * cd * argd));
                else if constexpr (F64 || QND_POT_ANGLE_FP64) return pot_factor(damp, cs * cd * argd);
                else return pot_factor(damp, csr * cr * argr);
            };
            CR xor_[QND_MAXD];
#pragma unroll
            for (int j = 0; j < QND_MAXD; ++j) xor_[j] = (CR)xo[j];
            const bool need_v = (fl & (PF_V1 | PF_V2 | PF_OBS | PF_SRC)) != 0;
            // Groups of four points: each operator is one unrolled loop over the
            // group with its mode tests outside, so the points are independent
            // chains the scheduler can interleave (a group bounds the registers
            // this takes).  Every point sees the operators in the same order
            // (potential, adjoint terms, V1, mask, observables, V2), and the
            // partial sums take the points in order m = 0 .. VPT-1.
            constexpr int G = VPT == 4 ? 4 : 2;  // points per group (registers vs interleaving)
#pragma unroll
            for (int m0 = 0; m0 < VPT; m0 += G) {
                C<CR> v[G];
                double xc[G], rr[G];
                CR xr[G], Vc[G];
                PR V[G], dVdu[G];
#pragma unroll
                for (int i = 0; i < G; ++i) {
                    const int m = m0 + i;
                    v[i] = scl<CR>(x[m], (CR)a.scale1);
                    xc[i] = coord_x(P, ax, t + m * T);
                    xr[i] = (CR)xc[i];
                    rr[i] = xc[i] * xc[i] + rro;
                    V[i] = (PR)0;
                    dVdu[i] = (PR)0;
                    Vc[i] = P.cap_on ? (CR)(cap_at(m) + Vco) : (CR)0;
                }
                // V from the table where only V is needed; V and dV/du from
                // the closed forms where the observables or the adjoint need
                // the derivative (the same FP64 arithmetic as the table's)
                const bool need_dv = (fl & (PF_OBS | PF_SRC | PF_GRAD)) != 0;
                if (VPT == 4 && a.vtab && !need_dv) {
#pragma unroll
                    for (int i = 0; i < G; ++i) V[i] = (PR)vtv[m0 + i];
                } else if (need_v) {
                    potential_pts<PR, G, !F64 && QND_POT_F64>(P, rr, V, dVdu);
                }
                if (fl & (PF_GRAD | PF_SRC)) {
#pragma unroll
                    for (int i = 0; i < G; ++i) {
                        const int m = m0 + i;
                        // adjoint: u = costate here; phi / y = forward state saved here
                        const C<CR> f = (KIND == PK_POT && (fl & PF_SRCM)) ? ysv[m]
                                                                           : cvtc<CR>(a.save[lbase + (int64_t)(t + m * T) * estride]);
                        double dV_da, dVdu_da;
                        softcore_derivs(P, rr[i], dV_da, dVdu_da);
                        if (fl & PF_GRAD) {
                            const double im = (double)v[i].x * (double)f.y - (double)v[i].y * (double)f.x;
                            obs[0] += dV_da * im;
#pragma unroll
                            for (int j = 0; j < QND_MAXD; ++j)
                                if (j < P.dim) obs[1 + j] += ((j == ax) ? xc[i] : xo[j]) * im;
                        } else {
                            double wsum = a.src_norm, dwsum = 0.0;
                            const double *srow = a.B.srcr ? a.B.srcr + (L / a.B.lpi) * a.B.srcr_stride : a.src;
#pragma unroll
                            for (int j = 0; j < QND_MAXD; ++j) {
                                if (j < P.dim) {
                                    const double xj = (j == ax) ? xc[i] : xo[j];
                                    wsum += srow[j] * (-2.0 * xj * dVdu[i] - fv[j]);
                                    dwsum += srow[j] * (-2.0 * xj * dVdu_da);
                                }
                            }
                            v[i] = mk<CR>((CR)((double)v[i].x + wsum * (double)f.x),
                                         (CR)((double)v[i].y + wsum * (double)f.y));
                            obs[0] += dwsum * ((double)f.x * (double)f.x + (double)f.y * (double)f.y);
                        }
                    }
                }
                if (fl & PF_V1) {
                    if (vmode == VM_IMAG) {
#pragma unroll
                        for (int i = 0; i < G; ++i) {
                            v[i] = scl<CR>(v[i], real_exp<CR>(-a.cv1 * V[i]));
                        }
                    } else if (vmode == VM_LENGTH) {
                        // without a CAP the damping argument is the constant 0 (no exponential)
                        if (P.cap_on) {
#pragma unroll
                            for (int i = 0; i < G; ++i)
                                v[i] = fac_mul<CR>(v[i], pf(-cv1 * Vc[i], cv1, a.cv1, V[i] + (xr[i] * fva + Wvr),
                                                           (double)V[i] + (xc[i] * fv[ax] + Wv)));
                        } else {
#pragma unroll
                            for (int i = 0; i < G; ++i)
                                v[i] = fac_mul<CR>(v[i], pf((CR)0, cv1, a.cv1, V[i] + (xr[i] * fva + Wvr),
                                                           (double)V[i] + (xc[i] * fv[ax] + Wv)));
                        }
                    } else {
                        if (P.cap_on) {
#pragma unroll
                            for (int i = 0; i < G; ++i)
                                v[i] = fac_mul<CR>(v[i], pf(-cv1 * Vc[i], cv1, a.cv1, V[i], (double)V[i]));
                        } else {
#pragma unroll
                            for (int i = 0; i < G; ++i)
                                v[i] = fac_mul<CR>(v[i], pf((CR)0, cv1, a.cv1, V[i], (double)V[i]));
                        }
                    }
                }
                if ((fl & PF_MASK) && !(fl & (PF_PREKX | PF_POSTKX)) && P.mask_on) {
#pragma unroll
                    for (int i = 0; i < G; ++i) {
                        const int m = m0 + i;
                        v[i] = scl<CR>(v[i], (CR)(msk_at(m) * mko));
                    }
                }
                if (fl & PF_OBS) {
#pragma unroll
                    for (int i = 0; i < G; ++i) {
                        const CR d = v[i].x * v[i].x + v[i].y * v[i].y;
                        add_obs(0, d);
#pragma unroll
                        for (int j = 0; j < QND_MAXD; ++j) {
                            if (j < P.dim) {
                                const CR xj = (j == ax) ? xr[i] : xor_[j];
                                add_obs(1 + j, d * xj);
                                add_obs(1 + QND_MAXD + j, d * dVdu[i] * (CR)2 * xj);
                            }
                        }
                    }
                }
                if (fl & PF_V2) {
                    if (vmode == VM_IMAG) {
#pragma unroll
                        for (int i = 0; i < G; ++i) {
                            v[i] = scl<CR>(v[i], real_exp<CR>(-a.cv2 * V[i]));
                        }
                    } else if (P.cap_on) {
#pragma unroll
                        for (int i = 0; i < G; ++i)
                            v[i] = fac_mul<CR>(v[i], pf(-cv2 * Vc[i], cv2, a.cv2, V[i], (double)V[i]));
                    } else {
#pragma unroll
                        for (int i = 0; i < G; ++i) v[i] = fac_mul<CR>(v[i], pf((CR)0, cv2, a.cv2, V[i], (double)V[i]));
                    }
                }
                if (fl & PF_SAVE) {
#pragma unroll
                    for (int i = 0; i < G; ++i) {
                        const int m = m0 + i;
                        a.save[lbase + (int64_t)(t + m * T) * estride] = cvtc<R>(v[i]);
                    }
                }
#pragma unroll
                for (int i = 0; i < G; ++i) {
                    const int m = m0 + i;
                    x[m] = v[i];
                }
            }
            if (!F64 && (fl & PF_OBS)) {
#pragma unroll
                for (int q = 0; q < QND_NOBS; ++q) obs[q] += (double)obr[q];
            }
            if (fl & (PF_OBS | PF_GRAD | PF_SRC))
                block_sum_vec<QND_NOBS>(obs, red, a.partials + (int64_t)tb * a.npart);
        }

        // ---- second transform ----
        if (KINK || (fl & PF_SECOND)) {
            transform<CR, -DIR1, VPT>(x, bufA, bufB, wc, n, log2n, T, t, tw, tw4, wave_line);
            const CR s2 = (CR)a.scale2;
            if ((KIND == PK_POT1D || KIND == PK_P1DI) && (fl & PF_KPOST)) {
#pragma unroll
                for (int m = 0; m < VPT; ++m) {
                    const double p = coord_p(P, ax, t + m * T);
                    const double K = .5 * p * p;
                    const C<CR> v = scl<CR>(x[m], s2);
                    if (kmode == KM_IMAG) x[m] = scl<CR>(v, real_exp<CR>(-a.ck_post * K));
                    else if (kmode == KM_VELOCITY)
                        x[m] = fac_mul<CR>(v, phase_fac<R, CR>(0.0, cs * a.ck_post * (K + (p + .5 * fv[ax]) * fv[ax])));
                    else x[m] = fac_mul<CR>(v, phase_fac<R, CR>(0.0, cs * a.ck_post * K));
                }
            } else if (s2 != (CR)1) {
#pragma unroll
                for (int m = 0; m < VPT; ++m) x[m] = scl<CR>(x[m], s2);
            }
        }
        if (KIND == PK_POT && (fl & PF_POSTKX)) {
            // adjoint of PF_PREKX: the conjugated axis-0 part of K(a_0), then
            // the axis-0 mask, on the state just returned to axis-0 Fourier space
#pragma unroll
            for (int m = 0; m < VPT; ++m) x[m] = fac_mul<CR>(x[m], tab_fac_x<CR>(atab_at(m), cs < 0.0));
            transform<CR, DIR1, VPT>(x, bufA, bufB, wc, n, log2n, T, t, tw, tw4, wave_line);
            const bool mk0 = (fl & PF_MASK) && P.mask_on;
#pragma unroll
            for (int m = 0; m < VPT; ++m) {
                const CR f = mk0 ? (CR)(msk_at(m) * a.scale1) : (CR)a.scale1;
                x[m] = scl<CR>(x[m], f);
            }
            transform<CR, -DIR1, VPT>(x, bufA, bufB, wc, n, log2n, T, t, tw, tw4, wave_line);
        }
        if (KIND == PK_KINS) {
            // merged step boundary: the state after the step's last kinetic
            // factor is the record's state (saved for the observable pass);
            // then this axis's part of the owed mask and of the next step's
            // first kinetic factor
            if (fl & PF_SAVE) {
#pragma unroll
                for (int m = 0; m < VPT; ++m)
                    a.save[lbase + (int64_t)(t + m * T) * estride] = cvtc<R>(r32 ? round_to_c64<CR>(x[m]) : x[m]);
            }
            if ((fl & PF_MASK) && P.mask_on) {
#pragma unroll
                for (int m = 0; m < VPT; ++m) x[m] = scl<CR>(x[m], (CR)kmask_at(m));
            }
            transform<CR, DIR1, VPT>(x, bufA, bufB, wc, n, log2n, T, t, tw, tw4, wave_line);
            const CR s2 = (CR)a.scale2;
#pragma unroll
            for (int m = 0; m < VPT; ++m)
                x[m] = scl<CR>(fac_mul<CR>(x[m], tab_fac_x<CR>(atab_at(m), cs < 0.0)), s2);
            transform<CR, -DIR1, VPT>(x, bufA, bufB, wc, n, log2n, T, t, tw, tw4, wave_line);
        }
    }

    // ---- store ----
    if (!(POTK && (fl & PF_NOSTORE))) {
        if (r32) {
#pragma unroll
            for (int m = 0; m < VPT; ++m) x[m] = round_to_c64<CR>(x[m]);
        }
#pragma unroll
        for (int m = 0; m < VPT; ++m) data[lbase + (int64_t)(t + m * T) * estride] = cvtc<R>(x[m]);
        if constexpr (X64) {
            // the dots below pair the state with stored values: use the stored one too
#pragma unroll
            for (int m = 0; m < VPT; ++m) x[m] = cvtc<CR>(cvtc<R>(x[m]));
        }
    }

    // ---- Gram-Schmidt dots on the stored representation: <prev_j|v> for
    // j < nprev, then <v|v>, in chunks of QND_GS_CHUNK.  The previous states'
    // values are requested QND_GS_G states at a time; each dot is summed over
    // the wavefront and parked in LDS, and one barrier per chunk precedes the
    // sums over the wavefronts (in wavefront order: deterministic). ----
    if (POTK && (fl & PF_GS_DOTS)) {
        __shared__ double gsred[2 * QND_GS_CHUNK * 16];
        constexpr int G = QND_GS_G(CR);
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        const int nw = (blockDim.x + 63) >> 6;
        const int nq = a.nprev + 1;
        for (int q0 = 0; q0 < nq; q0 += QND_GS_CHUNK) {
            const int q1 = nq < q0 + QND_GS_CHUNK ? nq : q0 + QND_GS_CHUNK;
            for (int j0 = q0; j0 < q1; j0 += G) {
                C<R> p[G][VPT];
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    if (j0 + g < q1 && j0 + g < a.nprev) {
                        const C<R> *pj = a.prev + (int64_t)(j0 + g) * a.slot_stride;
#pragma unroll
                        for (int m = 0; m < VPT; ++m) p[g][m] = pj[lbase + (int64_t)(t + m * T) * estride];
                    }
                }
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    const int j = j0 + g;
                    if (j < q1) {
                        double sr = 0.0, si = 0.0;
                        if (j < a.nprev) {
#pragma unroll
                            for (int m = 0; m < VPT; ++m) {
                                sr += (double)p[g][m].x * (double)x[m].x + (double)p[g][m].y * (double)x[m].y;
                                si += (double)p[g][m].x * (double)x[m].y - (double)p[g][m].y * (double)x[m].x;
                            }
                        } else {
#pragma unroll
                            for (int m = 0; m < VPT; ++m)
                                sr += (double)x[m].x * (double)x[m].x + (double)x[m].y * (double)x[m].y;
                        }
                        sr = wave_sum(sr);
                        si = wave_sum(si);
                        if (lane == 0) {
                            gsred[(2 * (j - q0)) * 16 + wv] = sr;
                            gsred[(2 * (j - q0) + 1) * 16 + wv] = si;
                        }
                    }
                }
            }
            lds_barrier();
            const int ncol = 2 * (q1 - q0);  // may exceed a small workgroup's threads
            for (int col = (int)threadIdx.x; col < ncol; col += (int)blockDim.x) {
                double sum = 0.0;
                for (int i = 0; i < nw; ++i) sum += gsred[col * 16 + i];
                a.partials[(int64_t)tb * a.npart + 2 * q0 + col] = sum;
            }
            lds_barrier();
        }
    }
}

template <typename R, int KIND, int DIR1, int VPT>
__global__ void __launch_bounds__(QND_PASS_THREADS)
    k_pass(C<R> *data, const typename TwT<R>::T *tw, int64_t S, unsigned geo, unsigned fl_arg, int nblk,
           const PassArgs<R> a)
{
    pass_body<R, KIND, DIR1, VPT, true>(data, tw, S, geo, fl_arg, nblk, a, (int)blockIdx.x);
}

// The same pass for a complex64 state held in FP64 registers (the complex64
// storage, FP64 arithmetic precision): the arguments and tables of the
// complex64 pass, each value widened at its load and rounded once at its
// store.  On gfx950, with packed FP32 arithmetic off, an FP64 FMA issues at
// the FP32 rate, so the transforms and factors cost no more VALU time and
// need none of the compensated FP32 products.
template <int KIND, int DIR1, int VPT>
__global__ void __launch_bounds__(QND_PASS_THREADS)
    k_pass_x(C<float> *data, const float4 *tw, int64_t S, unsigned geo, unsigned fl_arg, int nblk,
             const PassArgs<float> a)
{
    pass_body<float, KIND, DIR1, VPT, true, true>(data, tw, S, geo, fl_arg, nblk, a, (int)blockIdx.x);
}

// One launch of a pass over several grids ("classes": e.g. the instances of
// a sweep on 128^2, 256^2 and 384^2 grids).  Class c owns workgroups
// [blk0[c], blk0[c+1]) (blk0 multiples of 8, so each class keeps the
// XCD-aware tile order of its own pass): nblk[c] line tiles, then, for a
// finalizing kinetic pass, one finalize workgroup per instance, then idle
// padding.  Every class has the same workgroup size; its values per thread
// select the body.
#define QND_GROUP_MAX 4
template <typename R> struct GroupArgs {
    int ncls;
    int blk0[QND_GROUP_MAX + 1];
    int nblk[QND_GROUP_MAX];
    int vpt[QND_GROUP_MAX];
    unsigned geo[QND_GROUP_MAX];
    PassArgs<R> a[QND_GROUP_MAX];
};

// The descriptor lives in device memory (it is the same for every step of a
// call: the step counter moves the rows): indexing a kernel-argument array
// by the runtime class made the compiler copy the whole block to scratch.
template <typename R, int KIND, int DIR1>
__global__ void __launch_bounds__(QND_PASS_THREADS) k_pass_group(const GroupArgs<R> *__restrict__ gp)
{
    const GroupArgs<R> &g = *gp;
    const int w = (int)blockIdx.x;
    int c = 0;
#pragma unroll
    for (int k = 1; k < QND_GROUP_MAX; ++k)
        if (k < g.ncls && w >= g.blk0[k]) c = k;
    const int bid = w - g.blk0[c];
    const PassArgs<R> &a = g.a[c];
    const int extra = (KIND == PK_KIN && (a.flags & PF_FINALIZE)) ? a.B.nb : 0;
    const int nblk = g.nblk[c];
    if (bid >= nblk + extra) return;  // padding up to the next class
    const int vpt = g.vpt[c];
    const unsigned geo = g.geo[c];
    if (vpt == 4) pass_body<R, KIND, DIR1, 4, false>(a.data, a.tw, a.S, geo, a.flags, nblk, a, bid);
    else if (vpt == 12) pass_body<R, KIND, DIR1, 12, false>(a.data, a.tw, a.S, geo, a.flags, nblk, a, bid);
    else pass_body<R, KIND, DIR1, 8, false>(a.data, a.tw, a.S, geo, a.flags, nblk, a, bid);
}

// ---------------------------------------------------------------------
// Every line-pass kernel the library launches.  They are compiled in
// several kernel translation units (qnd_kernels_*.hip, one share each, in
// parallel) and declared extern in the host translation unit
// (qnd_api.hip), which then compiles no pass kernel itself.
// QND_PASS_SET_k(X, R) lists share k as X(R, KIND, DIR); the values per
// thread 4, 8 and 12 of each are one entry.
// ---------------------------------------------------------------------
#define QND_PASS_SET_0(X, R) X(R, PK_POTL, +1) X(R, PK_KINP, -1) X(R, PK_POT, +1) X(R, PK_PLAIN, -1) \
                             X(R, PK_PLAIN, +1)
#define QND_PASS_SET_1(X, R) X(R, PK_KIN, -1) X(R, PK_KINS, -1) X(R, PK_POTX, +1) X(R, PK_POTO, +1) \
                             X(R, PK_POTV, +1) X(R, PK_KINV, -1) X(R, PK_POTB, +1)
#define QND_PASS_SET_2(X, R) X(R, PK_POTI, +1) X(R, PK_KINI, -1) X(R, PK_P1DI, +1) X(R, PK_POT1D, +1) \
                             X(R, PK_KINR, -1) X(R, PK_POTS, +1) X(R, PK_POTG, +1) X(R, PK_POTJ, +1)
#define QND_GROUP_SET(X, R) X(R, PK_POT, +1) X(R, PK_POTL, +1) X(R, PK_KIN, -1) X(R, PK_KINP, -1)

#define QND_PASS_ARGS(R) C<R> *, const typename TwT<R>::T *, int64_t, unsigned, unsigned, int, const PassArgs<R>
#define QND_PASS_EXTERN(R, K, D)                                                                     \
    extern template __global__ void k_pass<R, K, D, 4>(QND_PASS_ARGS(R));                           \
    extern template __global__ void k_pass<R, K, D, 8>(QND_PASS_ARGS(R));                           \
    extern template __global__ void k_pass<R, K, D, 12>(QND_PASS_ARGS(R));
#define QND_PASS_INSTANTIATE(R, K, D)                                                                \
    template __global__ void k_pass<R, K, D, 4>(QND_PASS_ARGS(R));                                  \
    template __global__ void k_pass<R, K, D, 8>(QND_PASS_ARGS(R));                                  \
    template __global__ void k_pass<R, K, D, 12>(QND_PASS_ARGS(R));
#define QND_PASSX_ARGS C<float> *, const float4 *, int64_t, unsigned, unsigned, int, const PassArgs<float>
#define QND_PASSX_EXTERN(R, K, D)                                                                    \
    extern template __global__ void k_pass_x<K, D, 4>(QND_PASSX_ARGS);                              \
    extern template __global__ void k_pass_x<K, D, 8>(QND_PASSX_ARGS);                              \
    extern template __global__ void k_pass_x<K, D, 12>(QND_PASSX_ARGS);
#define QND_PASSX_INSTANTIATE(R, K, D)                                                               \
    template __global__ void k_pass_x<K, D, 4>(QND_PASSX_ARGS);                                     \
    template __global__ void k_pass_x<K, D, 8>(QND_PASSX_ARGS);                                     \
    template __global__ void k_pass_x<K, D, 12>(QND_PASSX_ARGS);
#define QND_GROUP_EXTERN(R, K, D) extern template __global__ void k_pass_group<R, K, D>(const GroupArgs<R> *);
#define QND_GROUP_INSTANTIATE(R, K, D) template __global__ void k_pass_group<R, K, D>(const GroupArgs<R> *);

#ifdef QND_HOST_TU
QND_PASS_SET_0(QND_PASS_EXTERN, float) QND_PASS_SET_0(QND_PASS_EXTERN, double)
QND_PASS_SET_1(QND_PASS_EXTERN, float) QND_PASS_SET_1(QND_PASS_EXTERN, double)
QND_PASS_SET_2(QND_PASS_EXTERN, float) QND_PASS_SET_2(QND_PASS_EXTERN, double)
QND_GROUP_SET(QND_GROUP_EXTERN, float) QND_GROUP_SET(QND_GROUP_EXTERN, double)
QND_PASS_SET_0(QND_PASSX_EXTERN, float) QND_PASS_SET_1(QND_PASSX_EXTERN, float)
QND_PASS_SET_2(QND_PASSX_EXTERN, float)
#endif

// advances the device step counter of a replayed step sequence
static __global__ void k_step_advance(int64_t *step)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

// ---------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------

// out[q] = scale * sum_w part[w*npart + q], one wavefront per q
static __global__ void __launch_bounds__(256) k_sum_partials(const double *__restrict__ part, int nwg, int npart,
                                                      int nq, double scale, double *out)
{
    const int wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    for (int q = blockIdx.x * nwv + wv; q < nq; q += gridDim.x * nwv) {
        const double s = wave_sum_column(part, nwg, npart, q);
        if ((threadIdx.x & 63) == 0) out[q] = s * scale;
    }
}

// out[p][q] = sum_w part[p][w][q] for npass blocks of [nwg][npart] partials,
// one wavefront per (p, q), fixed order
static __global__ void __launch_bounds__(256) k_sum_blocks(const double *__restrict__ part, int npass, int nwg, int npart,
                                                    double *out)
{
    const int wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    for (int pq = blockIdx.x * nwv + wv; pq < npass * npart; pq += gridDim.x * nwv) {
        const int p = pq / npart, q = pq - p * npart;
        const double s = wave_sum_column(part + (int64_t)p * nwg * npart, nwg, npart, q);
        if ((threadIdx.x & 63) == 0) out[pq] = s;
    }
}

// one block per instance
static __global__ void __launch_bounds__(1024) k_obs_finalize(const double *__restrict__ part, int nwg, int dim,
                                                       double dNr, ObsSpec spec, Batch B, double *rec)
{
    __shared__ double red[16 * QND_NOBS];
    obs_finalize_instance(part, nwg, dim, dNr, spec, B, blockIdx.x, rec, red);
}

// First-level sums of a [nrows][npart] partials array: block b sums the
// rows of its fixed slice [b * per, (b + 1) * per), per = ceil(nrows /
// gridDim.x), into out[b][0 .. npart).  The slices and the order inside one
// depend on nrows and the grid only, so the sums are deterministic.
static __global__ void __launch_bounds__(256) k_row_blocks(const double *__restrict__ part, int nrows, int npart,
                                                    double *out)
{
    __shared__ double red[16 * 8];
    const int per = (nrows + (int)gridDim.x - 1) / (int)gridDim.x;
    const int r0 = (int)blockIdx.x * per, r1 = min(nrows, r0 + per);
    const int n = r1 > r0 ? r1 - r0 : 0;
    for (int q0 = 0; q0 < npart; q0 += 8) {
        double tot[8];
        block_column_sums<8>(part + (int64_t)r0 * npart, n, npart, q0, tot, red);
        if (threadIdx.x == 0)
            for (int q = 0; q < 8 && q0 + q < npart; ++q) out[(int64_t)blockIdx.x * npart + q0 + q] = tot[q];
        lds_barrier();
    }
}

// Modified Gram-Schmidt coefficients (one thread) from the scaled dots in
// the stored representation (scale = dNr / n0 for the axis-0 Fourier
// representation): b_j = <prev_j|v>, nv = <v|v>;
// c_j = b_j - sum_{i<j} c_i G_{j i};
// |v'|^2 = nv - 2 Re sum_j c_j conj(b_j) + sum_{ij} conj(c_i) c_j G_{ij}.
// coef = [c_0 .. c_{N-1} (re,im), invnorm, flag, norm, step of the first collapse]
// `step` is the caller's index of the step whose dots these are: the first
// collapse (norm below 1e-12) records it beside the sticky flag, and the
// return value says whether this step collapsed.
__device__ __forceinline__ bool gs_coefficients(const double *b, int nprev, const double *__restrict__ gram,
                                                double *cre, double *cim, double *coef, double step)
{
    {
        const double nv = b[2 * nprev];
        for (int j = 0; j < nprev; ++j) {
            double r = b[2 * j], im = b[2 * j + 1];
            for (int i = 0; i < j; ++i) {
                const double gr = gram[2 * (j * nprev + i)], gi = gram[2 * (j * nprev + i) + 1];
                r -= cre[i] * gr - cim[i] * gi;
                im -= cre[i] * gi + cim[i] * gr;
            }
            cre[j] = r;
            cim[j] = im;
        }
        double n2 = nv;
        for (int j = 0; j < nprev; ++j) n2 -= 2.0 * (cre[j] * b[2 * j] + cim[j] * b[2 * j + 1]);
        for (int i = 0; i < nprev; ++i)
            for (int j = 0; j < nprev; ++j) {
                const double gr = gram[2 * (i * nprev + j)], gi = gram[2 * (i * nprev + j) + 1];
                const double qr = cre[i] * cre[j] + cim[i] * cim[j];
                const double qi = cre[i] * cim[j] - cim[i] * cre[j];
                n2 += qr * gr - qi * gi;
            }
        for (int j = 0; j < nprev; ++j) {
            coef[2 * j] = cre[j];
            coef[2 * j + 1] = cim[j];
        }
        if (n2 < 1.e-12) {  // sticky flag; the norm and step of the first collapse are kept
            if (coef[2 * nprev + 1] == 0.0) {
                coef[2 * nprev + 2] = n2;
                coef[2 * nprev + 3] = step;
            }
            coef[2 * nprev] = 0.0;
            coef[2 * nprev + 1] = 1.0;
            return true;
        }
        coef[2 * nprev] = 1.0 / sqrt(n2);
        if (coef[2 * nprev + 1] == 0.0) coef[2 * nprev + 2] = n2;
        return false;
    }
}

static __global__ void __launch_bounds__(1024) k_gs_coeffs(const double *__restrict__ part, int nwg, int npart,
                                                    int nprev, double scale, const double *__restrict__ gram,
                                                    double *coef, double step)
{
    __shared__ double b[2 * 65 + 8];
    __shared__ double cre[64], cim[64];
    __shared__ double red[16 * 8];
    // column sums of the dot partials, eight columns at a time over the whole block
    const int ncol = 2 * (nprev + 1);
    for (int q0 = 0; q0 < ncol; q0 += 8) {
        double tot[8];
        block_column_sums<8>(part, nwg, npart, q0, tot, red);
        if (threadIdx.x == 0)
            for (int q = 0; q < 8 && q0 + q < ncol; ++q) b[q0 + q] = tot[q] * scale;
    }
    lds_barrier();
    if (threadIdx.x == 0) (void)gs_coefficients(b, nprev, gram, cre, cim, coef, step);
}

// ---------------------------------------------------------------------
// flat elementwise and reduction kernels (representation-agnostic helpers
// for the per-state observables, Gram matrix and standalone updates)
// ---------------------------------------------------------------------
__device__ __forceinline__ void flat_index(const Phys &P, int64_t f, int *idx)
{
    if (P.pow2) {
        int sh = 0;
#pragma unroll
        for (int j = QND_MAXD - 1; j >= 0; --j) {
            idx[j] = 0;
            if (j < P.dim) {
                idx[j] = (int)((f >> sh) & (P.n[j] - 1));
                sh += P.lg[j];
            }
        }
        return;
    }
#pragma unroll
    for (int j = QND_MAXD - 1; j >= 0; --j) {
        idx[j] = 0;
        if (j < P.dim) {
            idx[j] = (int)(f % P.n[j]);
            f /= P.n[j];
        }
    }
}
// idx[j] for a runtime j without dynamic register indexing
// Kinetic factor tables: entry f of table q is e^{-i c_q K(f)}, K = sum_j
// p_j^2 / 2 over the row-major multi-index of f, with the terms in the
// kinetic pass's order (the last axis added last) and the factor formed in
// FP64.
template <typename T>
__global__ void __launch_bounds__(256) k_ktab(T *tab, const double *__restrict__ coef, int ncoef, int64_t N, Phys P)
{
    const int ax = P.dim - 1;
    for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < N; f += (int64_t)gridDim.x * blockDim.x) {
        int idx[QND_MAXD];
        flat_index(P, f, idx);
        double Ko = 0.0;
#pragma unroll
        for (int j = 0; j < QND_MAXD; ++j) {
            if (j < P.dim && j != ax) {
                const double pj = coord_p(P, j, idx[j]);
                Ko += .5 * pj * pj;
            }
        }
        const double p = coord_p(P, ax, idx[ax]);
        const double K = Ko + .5 * p * p;
        for (int q = 0; q < ncoef; ++q) {
            double sn, cn;
            sincos(coef[q] * K, &sn, &cn);
            tab_store(tab + (int64_t)q * N + f, cn, -sn);
        }
    }
}

// Potential table: V at every element of one instance, FP64, from the same
// closed forms (PotPoint<double>) as the passes.
static __global__ void __launch_bounds__(256) k_vtab(double *tab, int64_t N, Phys P)
{
    for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < N; f += (int64_t)gridDim.x * blockDim.x) {
        int idx[QND_MAXD];
        flat_index(P, f, idx);
        double rr = 0.0;
#pragma unroll
        for (int j = 0; j < QND_MAXD; ++j) {
            if (j < P.dim) {
                const double x = coord_x(P, j, idx[j]);
                rr += x * x;
            }
        }
        double dVdu;
        tab[f] = potential<double>(P, rr, dVdu);
    }
}

__device__ __forceinline__ int pick(const int *idx, int j)
{
    int v = idx[0];
#pragma unroll
    for (int k = 1; k < QND_MAXD; ++k) v = (j == k) ? idx[k] : v;
    return v;
}

enum ElemMode { EM_COPY = 0, EM_MUL_KIN = 1, EM_MUL_PSHIFT = 2, EM_GS_UPDATE = 3, EM_SCALE = 4, EM_RMASK = 5 };

template <typename R>
__global__ void k_elem(int mode, C<R> *dst, const C<R> *src, int axis, double shift, int64_t N,
                       Phys P, const C<R> *prev, int64_t slot_stride, int nprev, const double *coef)
{
    for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < N;
         f += (int64_t)gridDim.x * blockDim.x) {
        C<R> v = src[f];
        if (mode == EM_MUL_KIN) {
            int idx[QND_MAXD];
            flat_index(P, f, idx);
            double K = 0.0;
#pragma unroll
            for (int j = 0; j < QND_MAXD; ++j) {
                if (j < P.dim) {
                    const double p = coord_p(P, j, idx[j]);
                    K += .5 * p * p;
                }
            }
            v = scl<R>(v, (R)K);
        } else if (mode == EM_MUL_PSHIFT) {
            int idx[QND_MAXD];
            flat_index(P, f, idx);
            v = scl<R>(v, (R)(coord_p(P, axis, pick(idx, axis)) + shift));
        } else if (mode == EM_GS_UPDATE) {
            double vr = v.x, vi = v.y;
            for (int j = 0; j < nprev; ++j) {
                const C<R> pj = prev[(int64_t)j * slot_stride + f];
                const double cr = coef[2 * j], ci = coef[2 * j + 1];
                vr -= cr * (double)pj.x - ci * (double)pj.y;
                vi -= cr * (double)pj.y + ci * (double)pj.x;
            }
            const double inv = coef[2 * nprev];
            v = mk<R>((R)(vr * inv), (R)(vi * inv));
        } else if (mode == EM_SCALE) {
            v = scl<R>(v, (R)shift);
        } else if (mode == EM_RMASK) {
            // radial sin^2 mask of the final observables (src/io_functions.c:159-167);
            // R0 = shift, R1 = coef[0]
            int idx[QND_MAXD];
            flat_index(P, f, idx);
            double rr = 0.0;
#pragma unroll
            for (int j = 0; j < QND_MAXD; ++j) {
                if (j < P.dim) {
                    const double x = coord_x(P, j, idx[j]);
                    rr += x * x;
                }
            }
            const double rad = sqrt(rr), R0 = shift, R1 = coef[0];
            double fm = 0.0;
            if (rad >= R0 && rad <= R1) {
                const double sn = sin(.5 * 3.141592653589793 * (rad - R0) / (R1 - R0));
                fm = sn * sn;
            } else if (rad > R1) fm = 1.0;
            v = mk<R>((R)(fm * (double)v.x), (R)(fm * (double)v.y));
        }
        dst[f] = v;
    }
}

// Gram matrix of NP states (slots base + i * stride, i < NP) in one read of
// each: per-workgroup FP64 partials of sum conj(a_i) a_j for the pairs
// i <= j in row-major order of the upper triangle, (re, im) per pair, with
// the arithmetic of the RM_DOT reduction below.
#define QND_GRAM_MAX 8
template <typename R, int NP>
__global__ void __launch_bounds__(256) k_gram(const C<R> *__restrict__ base, int64_t stride, int64_t N,
                                              double *partials)
{
    constexpr int NPAIR = NP * (NP + 1) / 2;
    __shared__ double red[16 * 2 * NPAIR];
    double acc[2 * NPAIR];
#pragma unroll
    for (int q = 0; q < 2 * NPAIR; ++q) acc[q] = 0.0;
    for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < N;
         f += (int64_t)gridDim.x * blockDim.x) {
        double vr[NP], vi[NP];
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const C<R> v = base[i * stride + f];
            vr[i] = v.x;
            vi[i] = v.y;
        }
        int p = 0;
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int j = i; j < NP; ++j, ++p) {
                acc[2 * p] += vr[i] * vr[j] + vi[i] * vi[j];
                acc[2 * p + 1] += vr[i] * vi[j] - vi[i] * vr[j];
            }
    }
    block_reduce_lds(acc, 2 * NPAIR, red);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < 2 * NPAIR; ++q) partials[(int64_t)blockIdx.x * (2 * NPAIR) + q] = acc[q];
}

// Gram-Schmidt of a state v against NP previous states (slots base + j *
// stride) as two elementwise kernels over contiguous memory, for large grids
// (the passes' fused update and dots read the previous states with the
// pass's own tile shape, strided along axis 0).
// k_gs_dots: block b writes row b of `part`: the FP64 sums of Re and Im of
// conj(prev_j) v for j < NP, then |v|^2 and 0 -- the row layout of the fused
// dots, for k_gs_coeffs.
template <typename R, int NP>
__global__ void __launch_bounds__(256) k_gs_dots(const C<R> *__restrict__ v, const C<R> *__restrict__ base,
                                                 int64_t stride, int64_t N, double *part)
{
    constexpr int K = 2 * (NP + 1);
    __shared__ double red[16 * K];
    double acc[K];
#pragma unroll
    for (int q = 0; q < K; ++q) acc[q] = 0.0;
    for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < N;
         f += (int64_t)gridDim.x * blockDim.x) {
        const C<R> x = v[f];
        C<R> pj[NP > 0 ? NP : 1];
#pragma unroll
        for (int j = 0; j < NP; ++j) pj[j] = base[j * stride + f];
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            acc[2 * j] += (double)pj[j].x * (double)x.x + (double)pj[j].y * (double)x.y;
            acc[2 * j + 1] += (double)pj[j].x * (double)x.y - (double)pj[j].y * (double)x.x;
        }
        acc[2 * NP] += (double)x.x * (double)x.x + (double)x.y * (double)x.y;
    }
    block_reduce_lds(acc, K, red);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < K; ++q) part[(int64_t)blockIdx.x * K + q] = acc[q];
}
// k_gs_update: v = (v - sum_j c_j prev_j) * invnorm from the coefficients of
// k_gs_coeffs (coef[2j], coef[2j+1], coef[2 NP]), in the order and FP64
// arithmetic of the fused update.
template <typename R, int NP>
__global__ void __launch_bounds__(256) k_gs_update(C<R> *__restrict__ v, const C<R> *__restrict__ base,
                                                   int64_t stride, int64_t N, const double *__restrict__ coef)
{
    double cr[NP > 0 ? NP : 1], ci[NP > 0 ? NP : 1];
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        cr[j] = coef[2 * j];
        ci[j] = coef[2 * j + 1];
    }
    const double inv = coef[2 * NP];
    for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < N;
         f += (int64_t)gridDim.x * blockDim.x) {
        const C<R> x = v[f];
        C<R> pj[NP > 0 ? NP : 1];
#pragma unroll
        for (int j = 0; j < NP; ++j) pj[j] = base[j * stride + f];
        double vr = x.x, vi = x.y;
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            vr -= cr[j] * (double)pj[j].x - ci[j] * (double)pj[j].y;
            vi -= cr[j] * (double)pj[j].y + ci[j] * (double)pj[j].x;
        }
        v[f] = mk<R>((R)(vr * inv), (R)(vi * inv));
    }
}

enum RedMode { RM_DOT = 0, RM_DOTX = 1, RM_PMOM = 2, RM_OBS = 3, RM_HCOMB = 4 };

// Per-workgroup partial sums (FP64).  Layout per workgroup: npart doubles.
//  RM_DOT   : sum conj(a) b                      -> 2
//  RM_DOTX  : sum conj(a) x_axis b               -> 2
//  RM_PMOM  : sum |a|^2 (p_j + alpha_j), j<dim   -> dim
//  RM_OBS   : [ |a|^2, |a|^2 x_j, |a|^2 dV/dx_j ] -> 1 + 2*QND_MAXD
//  RM_HCOMB : b <- c * fs + V b ; sum conj(a) b  -> 2   (b = H psi, c = kinetic part)
template <typename R>
__global__ void __launch_bounds__(256) k_red(int mode, const C<R> *a, C<R> *b, const C<R> *c, int axis,
                                             double fs, int64_t N, Phys P, const double *alpha,
                                             double *partials, int npart)
{
    __shared__ double red[16];
    double acc[QND_NOBS];
#pragma unroll
    for (int q = 0; q < QND_NOBS; ++q) acc[q] = 0.0;
    for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < N;
         f += (int64_t)gridDim.x * blockDim.x) {
        const C<R> va = a[f];
        const double ar = va.x, ai = va.y;
        if (mode == RM_DOT || mode == RM_DOTX) {
            const C<R> vb = b[f];
            double w = 1.0;
            if (mode == RM_DOTX) {
                int idx[QND_MAXD];
                flat_index(P, f, idx);
                w = coord_x(P, axis, pick(idx, axis));
            }
            acc[0] += (ar * (double)vb.x + ai * (double)vb.y) * w;
            acc[1] += (ar * (double)vb.y - ai * (double)vb.x) * w;
        } else if (mode == RM_PMOM) {
            int idx[QND_MAXD];
            flat_index(P, f, idx);
            const double d = ar * ar + ai * ai;
#pragma unroll
            for (int j = 0; j < QND_MAXD; ++j)
                if (j < P.dim) acc[j] += d * (coord_p(P, j, idx[j]) + alpha[j]);
        } else {
            int idx[QND_MAXD];
            flat_index(P, f, idx);
            double rr = 0.0;
#pragma unroll
            for (int j = 0; j < QND_MAXD; ++j) {
                if (j < P.dim) {
                    const double x = coord_x(P, j, idx[j]);
                    rr += x * x;
                }
            }
            if (mode == RM_OBS) {
                double dVdu;
                (void)potential<double>(P, rr, dVdu);
                const double d = ar * ar + ai * ai;
                acc[0] += d;
#pragma unroll
                for (int j = 0; j < QND_MAXD; ++j) {
                    if (j < P.dim) {
                        const double x = coord_x(P, j, idx[j]);
                        acc[1 + j] += d * x;
                        acc[1 + QND_MAXD + j] += d * dVdu * 2.0 * x;
                    }
                }
            } else {  // RM_HCOMB
                double dVdu;
                const double V = potential<double>(P, rr, dVdu);
                const C<R> hk = c[f], hb = b[f];
                const double hr = (double)hk.x * fs + V * (double)hb.x;
                const double hi = (double)hk.y * fs + V * (double)hb.y;
                b[f] = mk<R>((R)hr, (R)hi);
                acc[0] += ar * hr + ai * hi;
                acc[1] += ar * hi - ai * hr;
            }
        }
    }
    const int nv = (mode == RM_OBS) ? QND_NOBS : (mode == RM_PMOM ? P.dim : 2);
#pragma unroll
    for (int q = 0; q < QND_NOBS; ++q) {
        if (q < nv) {
            const double s = block_sum(acc[q], red);
            if (threadIdx.x == 0) partials[(int64_t)blockIdx.x * npart + q] = s;
        }
    }
}

// Photoelectron energy distribution (src/io_functions.c:172-199) on the
// unitary p-space state: partial[w][j] = sum over this workgroup's points
// inside pr_max of |psi|^2 exp(-(K - j de)^2 / (2 sigma^2)) * nf.
template <typename R>
__global__ void __launch_bounds__(256) k_ped(const C<R> *psi, int64_t N, Phys P, const double *pr_max4,
                                             int Ne, double de, double sigma, double nf, int chunk,
                                             double *partials)
{
    __shared__ double dens[1024];
    __shared__ double kin[1024];
    const int64_t p0 = (int64_t)blockIdx.x * chunk;
    const int cnt = (int)((N - p0) < chunk ? (N - p0) : chunk);
    for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
        const int64_t f = p0 + q;
        int idx[QND_MAXD];
        flat_index(P, f, idx);
        bool inside = true;
        double K = 0.0;
#pragma unroll
        for (int j = 0; j < QND_MAXD; ++j) {
            if (j < P.dim) {
                const double p = coord_p(P, j, idx[j]);
                if (fabs(p) > pr_max4[j]) inside = false;
                K += .5 * p * p;
            }
        }
        const C<R> v = psi[f];
        dens[q] = inside ? ((double)v.x * (double)v.x + (double)v.y * (double)v.y) : 0.0;
        kin[q] = K;
    }
    lds_barrier();
    const double inv2s2 = 1.0 / (sigma * sigma);
    for (int j = threadIdx.x; j < Ne; j += blockDim.x) {
        double s = 0.0;
        const double Ej = j * de;
        for (int q = 0; q < cnt; ++q) {
            const double dk = kin[q] - Ej;
            if (dens[q] != 0.0) s += dens[q] * exp(-.5 * dk * dk * inv2s2) * nf;
        }
        partials[(int64_t)blockIdx.x * Ne + j] = s;
    }
}

// ---------------------------------------------------------------------
// 1-D imaginary time in one launch (src/integrators.c:135-163 and
// src/QuantumND_imag.c:182-185, 265-311).  A 1-D line fits one workgroup, so
// every pass of the multi-launch form would be a single workgroup waiting on
// its own launch, load and store.  Here one workgroup of 2^L / VPT threads
// keeps the 2^L-point line in registers for all nsteps steps: per stage the
// inverse transform, the
// potential factors, the forward transform and the kinetic factor, in the
// operator order and arithmetic of the 1-D stage pass (PK_P1DI); after the
// last potential half-step the Gram-Schmidt dots against the previous states
// (read from memory), the coefficients of gs_coefficients and the update.
// The state is read at entry and stored at exit, in the axis-0 Fourier
// representation of the slots.  A collapse (norm below 1e-12) stops the loop
// at that step, with the flag, norm and step in coef as from k_gs_coeffs.
// ---------------------------------------------------------------------
struct Imag1D {
    double as[20];        // as_scaled: a_k h, k = 0 .. 2s-1 (s <= 10)
    int s;
    int nprev;            // previous states for Gram-Schmidt; -1: none
    int ldsStride;
    double inv_n;         // 1 / n (the inverse transform's scale)
    double gs_scale;      // dNr / n (dots in the Fourier representation)
};

template <typename R, int VPT, int L>
static __global__ void __launch_bounds__(512) k_imag1d(C<R> *psi, const C<R> *__restrict__ prev,
                                                       int64_t slot_stride, const typename TwT<R>::T *tw,
                                                       const Phys P, const Imag1D A,
                                                       const double *__restrict__ gram, double *coef,
                                                       int64_t nsteps)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int NQ = 2 * (QND_GRAM_MAX + 1);
    __shared__ double red[NQ * 16];
    __shared__ double bq[NQ], cre[QND_GRAM_MAX], cim[QND_GRAM_MAX];
    __shared__ double upd[2 * QND_GRAM_MAX + 2];  // coefficients, invnorm, stop
    const int t = threadIdx.x, T = blockDim.x;
    C<R> *bufA = reinterpret_cast<C<R> *>(smem_raw);
    C<R> *bufB = bufA + A.ldsStride;
    int wc = 0;
    Tw4<R> tw4;
    if constexpr (VPT == 4) tw4_load<R>(tw4, tw, L, T, t);
    using TE = typename TwT<R>::T;

    C<R> x[VPT];
#pragma unroll
    for (int m = 0; m < VPT; ++m) x[m] = psi[t + m * T];

    // V (FP64, as in the stage pass: closed forms, FP64 in both precisions)
    // at this thread's positions
    double V[VPT];
    {
        double rr[VPT], dV[VPT];
#pragma unroll
        for (int m = 0; m < VPT; ++m) {
            const double xc = coord_x(P, 0, t + m * T);
            rr[m] = xc * xc;
        }
        potential_pts<double, VPT, !std::is_same<R, double>::value && QND_POT_F64>(P, rr, V, dV);
    }
    const R sc = (R)A.inv_n;
    const int np = A.nprev;
    const int lane = t & 63, wv = t >> 6, nw = (T + 63) >> 6;

    for (int64_t it = 0; it < nsteps; ++it) {
        for (int ia = 0; ia < A.s; ++ia) {
            fft_fixed<R, 1, VPT, L, TE>(x, bufA, bufB, wc, t, tw, tw4, false);
#pragma unroll
            for (int m = 0; m < VPT; ++m) {
                C<R> v = scl<R>(x[m], sc);
                if (ia > 0) v = scl<R>(v, real_exp<R>(-A.as[2 * ia - 1] * V[m]));
                x[m] = scl<R>(v, real_exp<R>(-A.as[2 * ia] * V[m]));
            }
            fft_fixed<R, -1, VPT, L, TE>(x, bufA, bufB, wc, t, tw, tw4, false);
            const double ck = A.as[2 * ia] + A.as[2 * ia + 1];
#pragma unroll
            for (int m = 0; m < VPT; ++m) {
                const double p = coord_p(P, 0, t + m * T);
                x[m] = scl<R>(x[m], real_exp<R>(-ck * (.5 * p * p)));
            }
        }
        // the last potential half-step
        fft_fixed<R, 1, VPT, L, TE>(x, bufA, bufB, wc, t, tw, tw4, false);
#pragma unroll
        for (int m = 0; m < VPT; ++m)
            x[m] = scl<R>(scl<R>(x[m], sc), real_exp<R>(-A.as[2 * A.s - 1] * V[m]));
        fft_fixed<R, -1, VPT, L, TE>(x, bufA, bufB, wc, t, tw, tw4, false);
        if (np < 0) continue;

        // Gram-Schmidt: dots <prev_j|v> for j < np, then <v|v> (FP64; each
        // summed over the wavefront and parked in LDS, then over the
        // wavefronts in order)
#pragma unroll 1
        for (int j = 0; j <= np; ++j) {
            double sr = 0.0, si = 0.0;
            if (j < np) {
                const C<R> *pj = prev + (int64_t)j * slot_stride;
#pragma unroll
                for (int m = 0; m < VPT; ++m) {
                    const C<R> p = pj[t + m * T];
                    sr += (double)p.x * (double)x[m].x + (double)p.y * (double)x[m].y;
                    si += (double)p.x * (double)x[m].y - (double)p.y * (double)x[m].x;
                }
            } else {
#pragma unroll
                for (int m = 0; m < VPT; ++m) sr += (double)x[m].x * (double)x[m].x + (double)x[m].y * (double)x[m].y;
            }
            sr = wave_sum(sr);
            si = wave_sum(si);
            if (lane == 0) {
                red[(2 * j) * 16 + wv] = sr;
                red[(2 * j + 1) * 16 + wv] = si;
            }
        }
        lds_barrier();
        if (t == 0) {
            for (int q = 0; q <= 2 * np + 1; ++q) {
                double sum = 0.0;
                for (int i = 0; i < nw; ++i) sum += red[q * 16 + i];
                bq[q] = sum * A.gs_scale;
            }
            const bool collapsed = gs_coefficients(bq, np, gram, cre, cim, coef, (double)it);
            for (int j = 0; j < np; ++j) {
                upd[2 * j] = cre[j];
                upd[2 * j + 1] = cim[j];
            }
            upd[2 * QND_GRAM_MAX] = coef[2 * np];
            upd[2 * QND_GRAM_MAX + 1] = collapsed ? 1.0 : 0.0;
        }
        lds_barrier();
        if (upd[2 * QND_GRAM_MAX + 1] != 0.0) break;  // uniform: the collapse step ends the call
        // v' = (v - sum_j c_j prev_j) * invnorm, subtracting in the order j = 0, 1, ...
        double vr[VPT], vi[VPT];
#pragma unroll
        for (int m = 0; m < VPT; ++m) {
            vr[m] = (double)x[m].x;
            vi[m] = (double)x[m].y;
        }
#pragma unroll 1
        for (int j = 0; j < np; ++j) {
            const C<R> *pj = prev + (int64_t)j * slot_stride;
            const double cr = upd[2 * j], ci = upd[2 * j + 1];
#pragma unroll
            for (int m = 0; m < VPT; ++m) {
                const C<R> p = pj[t + m * T];
                vr[m] -= cr * (double)p.x - ci * (double)p.y;
                vi[m] -= cr * (double)p.y + ci * (double)p.x;
            }
        }
        const double gsinv = upd[2 * QND_GRAM_MAX];
#pragma unroll
        for (int m = 0; m < VPT; ++m) x[m] = mk<R>((R)(vr[m] * gsinv), (R)(vi[m] * gsinv));
        lds_barrier();  // red and upd are rewritten in the next step
    }
#pragma unroll
    for (int m = 0; m < VPT; ++m) psi[t + m * T] = x[m];
}

}  // namespace qnd
