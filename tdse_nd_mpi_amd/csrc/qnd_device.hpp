// qnd_device.hpp -- device code of libqnd for gfx950 (CDNA4, wave64).
//
// The split-operator step is built from "line passes": a workgroup loads a
// tile of TW lines of length n along one axis into LDS, runs an in-place
// Stockham FFT on them (radix 8/4/2, one butterfly set per thread, twiddles
// from an FP64/FP32 table), applies the fused pointwise operators of the
// step in registers, optionally transforms back, reduces observables in FP64
// and writes the tile back.  Only the state crosses HBM; potential, CAP,
// kinetic, laser-coupling and mask factors are evaluated analytically.
//
// Pass families (see qnd_api.hip for the schedules):
//   axis-0 pass ("potential pass"): IFFT_0 -> r-space ops -> FFT_0
//   last-axis pass ("kinetic pass"): FFT_{d-1} -> k-space ops -> IFFT_{d-1}
//   plain pass: one transform along an axis (representation changes)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qnd {

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
template <typename R> __device__ __forceinline__ C<R> cmulconj(C<R> a, C<R> b)  // a * conj(b)
{
    return mk<R>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
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

template <typename R, int DIR> __device__ __forceinline__ void dft8(C<R> *x)
{
    const R h = (R)0.70710678118654752440084436210484903928;
    C<R> e[4] = {x[0], x[2], x[4], x[6]};
    C<R> o[4] = {x[1], x[3], x[5], x[7]};
    dft4<R, DIR>(e);
    dft4<R, DIR>(o);
    // twiddles w^q, w = exp(DIR i pi/4)
    C<R> o1 = mk<R>(h * (o[1].x - (R)DIR * o[1].y), h * (o[1].y + (R)DIR * o[1].x));
    C<R> o2 = mul_di<R, DIR>(o[2]);
    C<R> o3 = mk<R>(h * (-o[3].x - (R)DIR * o[3].y), h * (-o[3].y + (R)DIR * o[3].x));
    x[0] = add<R>(e[0], o[0]);
    x[4] = sub<R>(e[0], o[0]);
    x[1] = add<R>(e[1], o1);
    x[5] = sub<R>(e[1], o1);
    x[2] = add<R>(e[2], o2);
    x[6] = sub<R>(e[2], o2);
    x[3] = add<R>(e[3], o3);
    x[7] = sub<R>(e[3], o3);
}

template <typename R, int RAD, int DIR> __device__ __forceinline__ void dftR(C<R> *x)
{
    if (RAD == 2) dft2<R, DIR>(x);
    else if (RAD == 4) dft4<R, DIR>(x);
    else dft8<R, DIR>(x);
}

// ---------------------------------------------------------------------
// Stockham radix stage on one line held in LDS.  The thread owns butterflies
// j = t + b*T (b < 8/RAD); every thread holds exactly 8 values.
//   inputs  x[r] = line[j + r*n/RAD]
//   twiddle x[r] *= W_n^{r k n/(Ns RAD)}, k = j mod Ns
//   outputs line[(j-k)*RAD + k + r*Ns] = DFT_RAD(x)[r]
// ---------------------------------------------------------------------
template <typename R, int RAD, int DIR>
__device__ __forceinline__ void stockham_stage(C<R> *line, int n, int T, int t, int Ns,
                                               const C<R> *__restrict__ tw, bool active)
{
    constexpr int NB = 8 / RAD;
    C<R> x[8];
    const int nR = n / RAD;
    if (active) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int j = t + b * T;
#pragma unroll
            for (int r = 0; r < RAD; ++r) x[b * RAD + r] = line[j + r * nR];
        }
    }
    __syncthreads();
    if (active) {
        const int stride = n / (Ns * RAD);
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int j = t + b * T;
            const int k = j & (Ns - 1);
            C<R> *xb = x + b * RAD;
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < RAD; ++r) {
                    C<R> w = tw[r * k * stride];
                    if (DIR > 0) w.y = -w.y;
                    xb[r] = mul<R>(xb[r], w);
                }
            }
            dftR<R, RAD, DIR>(xb);
            const int o = (j - k) * RAD + k;
#pragma unroll
            for (int r = 0; r < RAD; ++r) line[o + r * Ns] = xb[r];
        }
    }
    __syncthreads();
}

// Full transform of one line (n = 2^m >= 8): radix-8 stages, then a final
// radix-4 or radix-2 stage for the remainder.
template <typename R, int DIR>
__device__ __forceinline__ void fft_line(C<R> *line, int n, int log2n, int T, int t,
                                         const C<R> *__restrict__ tw, bool active)
{
    int Ns = 1;
    const int n8 = log2n / 3, rem = log2n - 3 * n8;
    for (int s = 0; s < n8; ++s) {
        stockham_stage<R, 8, DIR>(line, n, T, t, Ns, tw, active);
        Ns *= 8;
    }
    if (rem == 2) stockham_stage<R, 4, DIR>(line, n, T, t, Ns, tw, active);
    else if (rem == 1) stockham_stage<R, 2, DIR>(line, n, T, t, Ns, tw, active);
}

// ---------------------------------------------------------------------
// physics parameters (by value in every launch)
// ---------------------------------------------------------------------
#define QND_MAXD 4
struct Phys {
    int dim;
    int n[QND_MAXD];
    double dr[QND_MAXD];
    double L[QND_MAXD];   // n*dr
    double dp[QND_MAXD];  // 2 pi / L
    int pot_kind;         // 0 soft-core, 1 gaussian (Ne_short_2D)
    double pot_b, pot_c2;
    double cap0;
    double capL[QND_MAXD];  // edge where the CAP starts: n dr/2 - width
    int mask_on;
    double mask_x0[QND_MAXD], mask_w[QND_MAXD];
    double dNr;
};

__device__ __forceinline__ double coord_x(const Phys &P, int j, int i) { return -.5 * P.L[j] + i * P.dr[j]; }
__device__ __forceinline__ double coord_p(const Phys &P, int j, int i)
{
    return (i < P.n[j] / 2) ? i * P.dp[j] : (i - P.n[j]) * P.dp[j];
}
__device__ __forceinline__ double cap_axis(const Phys &P, int j, double x)
{
    double ax = fabs(x);
    return (ax >= P.capL[j]) ? P.cap0 * (ax - P.capL[j]) * (ax - P.capL[j]) : 0.0;
}
__device__ __forceinline__ double mask_axis(const Phys &P, int j, double x)
{
    double ax = fabs(x);
    if (ax <= P.mask_x0[j]) return 1.0;
    return pow(cos(1.5707963267948966 * (ax - P.mask_x0[j]) / P.mask_w[j]), 0.125);
}
// V(r) and dV/du (u = r^2) for the named potentials (src/meshgrid.c:140-153)
__device__ __forceinline__ double potential(const Phys &P, double rr, double *dVdu)
{
    double rad = sqrt(rr);
    double u = rad * rad;
    if (P.pot_kind == 0) {
        double den = sqrt(u + P.pot_c2);
        double g = (P.pot_b != 0.0) ? P.pot_b * exp(-u) : 0.0;
        if (dVdu) *dVdu = g / den + .5 * (1. + g) / (den * (u + P.pot_c2));
        return -(1. + g) / den;
    }
    double g = 2.7459 * exp(-.5 * rad * rad / 1.96);
    if (dVdu) *dVdu = g * .5 / 1.96;
    return -g;
}

#define QND_MAX_EXP_ARG 700.0     // src/integrators.h:12
#define QND_MAX_EXP_ARG_F32 80.0  // FP32 path: e^80 stays finite in float

template <typename R> struct Prec;
template <> struct Prec<double> { static constexpr double clamp = QND_MAX_EXP_ARG; };
template <> struct Prec<float> { static constexpr double clamp = QND_MAX_EXP_ARG_F32; };

// e^{arg} (cos ph - i sin ph): FP64 for complex128; for complex64 the angle is
// reduced mod 2 pi in FP64 and sin/cos are evaluated in FP32.
template <typename R> __device__ __forceinline__ C<R> phase_factor(double arg, double ph);
template <> __device__ __forceinline__ C<double> phase_factor<double>(double arg, double ph)
{
    double s, c;
    sincos(ph, &s, &c);
    double e = (arg == 0.0) ? 1.0 : exp(arg > QND_MAX_EXP_ARG ? QND_MAX_EXP_ARG : arg);
    return mk<double>(e * c, -e * s);
}
template <> __device__ __forceinline__ C<float> phase_factor<float>(double arg, double ph)
{
    const double twopi = 6.283185307179586476925286766559;
    double red = ph - twopi * rint(ph * (1.0 / twopi));
    float s, c;
    sincosf((float)red, &s, &c);
    float e = (arg == 0.0) ? 1.0f : (float)exp(arg > QND_MAX_EXP_ARG_F32 ? QND_MAX_EXP_ARG_F32 : arg);
    return mk<float>(e * c, -e * s);
}
template <typename R> __device__ __forceinline__ R real_exp(double arg)
{
    const double cl = Prec<R>::clamp;
    return (R)exp(arg > cl ? cl : arg);
}

// ---------------------------------------------------------------------
// wavefront / workgroup FP64 reductions (deterministic butterfly order)
// ---------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// Reduce v over the block; thread 0 gets the result.  red = LDS scratch of
// at least 16 doubles.  All threads must call.
__device__ __forceinline__ double block_sum(double v, double *red)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < nw; ++i) s += red[i];
    return s;
}

// ---------------------------------------------------------------------
// line pass
// ---------------------------------------------------------------------
enum PassFlags : unsigned {
    PF_GS_UPDATE = 1u << 0,  // prologue: v = (v - sum_j c_j prev_j) * invnorm
    PF_KPRE = 1u << 1,       // (d = 1) kinetic factor before the first transform
    PF_V1 = 1u << 2,         // potential factor #1 (after the first transform)
    PF_MASK = 1u << 3,       // cos^(1/8) mask
    PF_OBS = 1u << 4,        // r-space observable partial sums
    PF_V2 = 1u << 5,         // potential factor #2
    PF_SECOND = 1u << 6,     // second (reverse) transform
    PF_KPOST = 1u << 7,      // (d = 1) kinetic factor after the second transform
    PF_GS_DOTS = 1u << 8,    // Gram-Schmidt partial dots after the second transform
    PF_KMID = 1u << 9,       // kinetic factor after the first transform
};
enum KMode { KM_PHASE = 0, KM_VELOCITY = 1, KM_IMAG = 2 };
enum VMode { VM_LENGTH = 0, VM_VELOCITY = 1, VM_IMAG = 2 };

#define QND_NOBS (1 + 2 * QND_MAXD)

template <typename R> struct PassArgs {
    C<R> *data;
    const C<R> *tw;
    int axis, n, log2n, T, TW, ldsStride;
    int64_t S, nlines;
    unsigned flags;
    int dir_first;           // -1: forward first, +1: inverse first
    double scale1, scale2;   // after the first / second transform
    int kmode;
    double ck_pre, ck_mid, ck_post;
    double Ak[QND_MAXD];
    int vmode;
    double cv1, cv2;
    double Ev[QND_MAXD];
    // Gram-Schmidt
    const C<R> *prev;
    int64_t slot_stride;
    int nprev;
    const double *gs_coef;   // [2*nprev] c_j, then invnorm, then error flag
    double *partials;        // per-workgroup outputs
    int npart;
    Phys P;
};

// Line multi-index helpers: a line is identified by L in [0, nlines):
// o = L / S enumerates the axes before `axis`, L % S the axes after it.
__device__ __forceinline__ void line_index(const Phys &P, int axis, int64_t S, int64_t L, int *idx)
{
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
// idx[j] for a runtime j without dynamic register indexing
__device__ __forceinline__ int pick(const int *idx, int j)
{
    int v = idx[0];
#pragma unroll
    for (int k = 1; k < QND_MAXD; ++k) v = (j == k) ? idx[k] : v;
    return v;
}

template <typename R, int DIR1>
__device__ __forceinline__ void pass_body(const PassArgs<R> &a, C<R> *lds, double *red)
{
    const int tid = threadIdx.x;
    const int n = a.n, T = a.T, TW = a.TW, ldsS = a.ldsStride;
    const int64_t line0 = (int64_t)blockIdx.x * TW;
    const int nl = (int)((a.nlines - line0) < TW ? (a.nlines - line0) : TW);
    const Phys &P = a.P;
    const unsigned fl = a.flags;

    // ---- load the tile (with the Gram-Schmidt update prologue) ----
    const bool contig = (a.S == 1);
    int64_t base;
    if (contig) base = line0 * n;
    else {
        int64_t o = line0 / a.S;
        base = o * n * a.S + (line0 - o * a.S);
    }
    const int tot = nl * n;
    double gsinv = 1.0;
    if (fl & PF_GS_UPDATE) gsinv = a.gs_coef[2 * a.nprev];
    for (int idx = tid; idx < tot; idx += blockDim.x) {
        int l, i;
        int64_t g;
        if (contig) { l = idx / n; i = idx - l * n; g = base + idx; }
        else { i = idx / nl; l = idx - i * nl; g = base + (int64_t)i * a.S + l; }
        C<R> v = a.data[g];
        if (fl & PF_GS_UPDATE) {
            C<double> vd = mk<double>((double)v.x, (double)v.y);
            for (int j = 0; j < a.nprev; ++j) {
                C<R> pj = a.prev[(int64_t)j * a.slot_stride + g];
                double cr = a.gs_coef[2 * j], ci = a.gs_coef[2 * j + 1];
                vd.x -= cr * (double)pj.x - ci * (double)pj.y;
                vd.y -= cr * (double)pj.y + ci * (double)pj.x;
            }
            v = mk<R>((R)(vd.x * gsinv), (R)(vd.y * gsinv));
        }
        lds[l * ldsS + i] = v;
    }
    __syncthreads();

    // ---- thread <-> (line, t) ----
    const int l = tid / T, t = tid - (tid / T) * T;
    const bool active = l < nl;
    C<R> *line = lds + l * ldsS;
    int idx[QND_MAXD] = {0, 0, 0, 0};
    if (active) line_index(P, a.axis, a.S, line0 + l, idx);
    const int ax = a.axis;

    // per-line partial sums over the other axes (axis order as the reference)
    double Ko = 0.0, Wk = 0.0, rro = 0.0, Vco = 0.0, Wv = 0.0, mko = 1.0;
    double xo[QND_MAXD] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < QND_MAXD; ++j) {
        if (j >= P.dim || j == ax) continue;
        if (fl & (PF_KMID)) {
            double pj = coord_p(P, j, idx[j]);
            Ko += .5 * pj * pj;
            if (a.kmode == KM_VELOCITY) Wk += (pj + .5 * a.Ak[j]) * a.Ak[j];
        }
        if (fl & (PF_V1 | PF_V2 | PF_MASK | PF_OBS)) {
            double xj = coord_x(P, j, idx[j]);
            xo[j] = xj;
            rro += xj * xj;
            Vco += cap_axis(P, j, xj);
            Wv += xj * a.Ev[j];
            if ((fl & PF_MASK) && P.mask_on) mko *= mask_axis(P, j, xj);
        }
    }

    // ---- kinetic factor before the first transform (d == 1) ----
    if ((fl & PF_KPRE) && active) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int i = t + e * T;
            double p = coord_p(P, ax, i);
            double K = .5 * p * p;
            C<R> f;
            if (a.kmode == KM_IMAG) { R r = real_exp<R>(-a.ck_pre * K); f = mk<R>(r, 0); }
            else f = phase_factor<R>(0.0, a.ck_pre * K);
            line[i] = mul<R>(line[i], f);
        }
    }
    if (fl & PF_KPRE) __syncthreads();

    // ---- first transform ----
    fft_line<R, DIR1>(line, n, a.log2n, T, t, a.tw, active);

    // ---- middle: scale, kinetic or potential operators, observables ----
    double obs[QND_NOBS];
#pragma unroll
    for (int q = 0; q < QND_NOBS; ++q) obs[q] = 0.0;
    if (active) {
        const R s1 = (R)a.scale1;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int i = t + e * T;
            C<R> v = line[i];
            if (s1 != (R)1) v = scl<R>(v, s1);
            if (fl & PF_KMID) {
                double p = coord_p(P, ax, i);
                double K = Ko + .5 * p * p;
                C<R> f;
                if (a.kmode == KM_IMAG) { R r = real_exp<R>(-a.ck_mid * K); f = mk<R>(r, 0); }
                else if (a.kmode == KM_VELOCITY) {
                    double W = Wk + (p + .5 * a.Ak[ax]) * a.Ak[ax];
                    f = phase_factor<R>(0.0, a.ck_mid * (K + W));
                } else f = phase_factor<R>(0.0, a.ck_mid * K);
                v = mul<R>(v, f);
            }
            if (fl & (PF_V1 | PF_V2 | PF_MASK | PF_OBS)) {
                const double x = coord_x(P, ax, i);
                const double rr = x * x + rro;
                double dVdu = 0.0;
                const double V = potential(P, rr, (fl & PF_OBS) ? &dVdu : nullptr);
                const double Vc = cap_axis(P, ax, x) + Vco;
                if (fl & PF_V1) {
                    C<R> f;
                    if (a.vmode == VM_IMAG) { R r = real_exp<R>(-a.cv1 * V); f = mk<R>(r, 0); }
                    else if (a.vmode == VM_LENGTH) {
                        double W = x * a.Ev[ax] + Wv;
                        f = phase_factor<R>(-a.cv1 * Vc, a.cv1 * (V + W));
                    } else f = phase_factor<R>(-a.cv1 * Vc, a.cv1 * V);
                    v = mul<R>(v, f);
                }
                if ((fl & PF_MASK) && P.mask_on) {
                    R m = (R)(mask_axis(P, ax, x) * mko);
                    v = scl<R>(v, m);
                }
                if (fl & PF_OBS) {
                    double d = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
                    obs[0] += d;
#pragma unroll
                    for (int j = 0; j < QND_MAXD; ++j) {
                        if (j < P.dim) {
                            double xj = (j == ax) ? x : xo[j];
                            obs[1 + j] += d * xj;
                            obs[1 + QND_MAXD + j] += d * dVdu * 2.0 * xj;
                        }
                    }
                }
                if (fl & PF_V2) {
                    C<R> f;
                    if (a.vmode == VM_IMAG) { R r = real_exp<R>(-a.cv2 * V); f = mk<R>(r, 0); }
                    else f = phase_factor<R>(-a.cv2 * Vc, a.cv2 * V);
                    v = mul<R>(v, f);
                }
            }
            line[i] = v;
        }
    }
    if (fl & PF_OBS) {
        for (int q = 0; q < 1 + 2 * QND_MAXD; ++q) {
            double s = block_sum(obs[q], red);
            if (tid == 0) a.partials[(int64_t)blockIdx.x * a.npart + q] = s;
        }
    }
    __syncthreads();

    // ---- second transform ----
    if (fl & PF_SECOND) {
        fft_line<R, -DIR1>(line, n, a.log2n, T, t, a.tw, active);
        if (active) {
            const R s2 = (R)a.scale2;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int i = t + e * T;
                C<R> v = line[i];
                if (s2 != (R)1) v = scl<R>(v, s2);
                if (fl & PF_KPOST) {
                    double p = coord_p(P, ax, i);
                    double K = .5 * p * p;
                    C<R> f;
                    if (a.kmode == KM_IMAG) { R r = real_exp<R>(-a.ck_post * K); f = mk<R>(r, 0); }
                    else if (a.kmode == KM_VELOCITY) {
                        double W = (p + .5 * a.Ak[ax]) * a.Ak[ax];
                        f = phase_factor<R>(0.0, a.ck_post * (K + W));
                    } else f = phase_factor<R>(0.0, a.ck_post * K);
                    v = mul<R>(v, f);
                }
                line[i] = v;
            }
        }
        __syncthreads();
    }

    // ---- store (and Gram-Schmidt dots on the stored representation) ----
    for (int idx2 = tid; idx2 < tot; idx2 += blockDim.x) {
        int ll, i;
        int64_t g;
        if (contig) { ll = idx2 / n; i = idx2 - ll * n; g = base + idx2; }
        else { i = idx2 / nl; ll = idx2 - i * nl; g = base + (int64_t)i * a.S + ll; }
        a.data[g] = lds[ll * ldsS + i];
    }
    if (fl & PF_GS_DOTS) {
        // <prev_j | v> partials and |v|^2 over this tile (same element order
        // for every j: deterministic)
        for (int j = 0; j <= a.nprev; ++j) {
            double sr = 0.0, si = 0.0;
            for (int idx2 = tid; idx2 < tot; idx2 += blockDim.x) {
                int ll, i;
                int64_t g;
                if (contig) { ll = idx2 / n; i = idx2 - ll * n; g = base + idx2; }
                else { i = idx2 / nl; ll = idx2 - i * nl; g = base + (int64_t)i * a.S + ll; }
                C<R> v = lds[ll * ldsS + i];
                if (j < a.nprev) {
                    C<R> pj = a.prev[(int64_t)j * a.slot_stride + g];
                    sr += (double)pj.x * (double)v.x + (double)pj.y * (double)v.y;
                    si += (double)pj.x * (double)v.y - (double)pj.y * (double)v.x;
                } else {
                    sr += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
                }
            }
            double tr = block_sum(sr, red);
            double ti = block_sum(si, red);
            if (tid == 0) {
                a.partials[(int64_t)blockIdx.x * a.npart + 2 * j] = tr;
                a.partials[(int64_t)blockIdx.x * a.npart + 2 * j + 1] = ti;
            }
        }
    }
}

template <typename R>
__global__ void __launch_bounds__(1024) k_pass(PassArgs<R> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[16];
    C<R> *lds = reinterpret_cast<C<R> *>(smem_raw);
    if (a.dir_first < 0) pass_body<R, -1>(a, lds, red);
    else pass_body<R, +1>(a, lds, red);
}

// ---------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------

// Sum per-workgroup partials: out[q] = scale * sum_w part[w*npart + q]
// (one thread per q, sequential over w: deterministic).
__global__ void k_sum_partials(const double *__restrict__ part, int nwg, int npart, int nq,
                               double scale, double *out)
{
    for (int q = threadIdx.x; q < nq; q += blockDim.x) {
        double s = 0.0;
        for (int w = 0; w < nwg; ++w) s += part[(int64_t)w * npart + q];
        out[q] = s * scale;
    }
}

// Observables from r-space partial sums [n0, d_j, g_j] of one step:
// writes rec[0..width) per the code list (norm, dipole_r, dipole_a).
struct ObsSpec {
    int ncodes;
    int codes[8];
    double E[QND_MAXD];
    double dNr;
};
__global__ void k_obs_finalize(const double *__restrict__ part, int nwg, int npart, int dim,
                               ObsSpec spec, double *rec)
{
    __shared__ double tot[QND_NOBS];
    for (int q = threadIdx.x; q < QND_NOBS; q += blockDim.x) {
        double s = 0.0;
        for (int w = 0; w < nwg; ++w) s += part[(int64_t)w * npart + q];
        tot[q] = s * spec.dNr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int off = 0;
        for (int k = 0; k < spec.ncodes; ++k) {
            int c = spec.codes[k];
            if (c == 5) { rec[off++] = tot[0]; }                 // norm
            else if (c == 3) { for (int j = 0; j < dim; ++j) rec[off++] = tot[1 + j]; }
            else if (c == 6) { for (int j = 0; j < dim; ++j) rec[off++] = -tot[1 + QND_MAXD + j] - spec.E[j] * tot[0]; }
            else if (c == 1 || c == 2) { off += 1; }             // filled by the host path
            else if (c == 4) { off += dim; }
        }
    }
}

// Modified Gram-Schmidt coefficients from partial dots in the stored
// representation (scale = dNr / n0 for the axis-0 Fourier representation):
// b_j = <prev_j|v>, nv = <v|v>; c_j = b_j - sum_{i<j} c_i G_{j i};
// |v'|^2 = nv - 2 Re sum_j c_j conj(b_j) + sum_{ij} conj(c_i) c_j G_{ij}.
// coef = [c_0 .. c_{N-1} (re,im), invnorm, flag, norm]
__global__ void k_gs_coeffs(const double *__restrict__ part, int nwg, int npart, int nprev,
                            double scale, const double *__restrict__ gram, double *coef)
{
    __shared__ double b[2 * 65];
    for (int q = threadIdx.x; q < 2 * (nprev + 1); q += blockDim.x) {
        double s = 0.0;
        for (int w = 0; w < nwg; ++w) s += part[(int64_t)w * npart + q];
        b[q] = s * scale;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double nv = b[2 * nprev];
        double cr[64], ci[64];
        for (int j = 0; j < nprev; ++j) {
            double r = b[2 * j], im = b[2 * j + 1];
            for (int i = 0; i < j; ++i) {
                // G_{j i} = <prev_j|prev_i>
                double gr = gram[2 * (j * nprev + i)], gi = gram[2 * (j * nprev + i) + 1];
                r -= cr[i] * gr - ci[i] * gi;
                im -= cr[i] * gi + ci[i] * gr;
            }
            cr[j] = r;
            ci[j] = im;
        }
        double n2 = nv;
        for (int j = 0; j < nprev; ++j) n2 -= 2.0 * (cr[j] * b[2 * j] + ci[j] * b[2 * j + 1]);
        for (int i = 0; i < nprev; ++i)
            for (int j = 0; j < nprev; ++j) {
                double gr = gram[2 * (i * nprev + j)], gi = gram[2 * (i * nprev + j) + 1];
                // conj(c_i) c_j G_ij
                double qr = cr[i] * cr[j] + ci[i] * ci[j];
                double qi = cr[i] * ci[j] - ci[i] * cr[j];
                n2 += qr * gr - qi * gi;
            }
        for (int j = 0; j < nprev; ++j) { coef[2 * j] = cr[j]; coef[2 * j + 1] = ci[j]; }
        if (n2 < 1.e-12) { coef[2 * nprev] = 0.0; coef[2 * nprev + 1] = 1.0; }  // sticky flag
        else { coef[2 * nprev] = 1.0 / sqrt(n2); }
        coef[2 * nprev + 2] = n2;
    }
}

// ---------------------------------------------------------------------
// flat elementwise and reduction kernels (representation-agnostic helpers
// for the per-state observables, Gram matrix and standalone updates)
// ---------------------------------------------------------------------
__device__ __forceinline__ void flat_index(const Phys &P, int64_t f, int *idx)
{
#pragma unroll
    for (int j = QND_MAXD - 1; j >= 0; --j) {
        idx[j] = 0;
        if (j < P.dim) {
            idx[j] = (int)(f % P.n[j]);
            f /= P.n[j];
        }
    }
}

enum ElemMode { EM_COPY = 0, EM_MUL_KIN = 1, EM_MUL_PSHIFT = 2, EM_GS_UPDATE = 3, EM_SCALE = 4 };

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
                    double p = coord_p(P, j, idx[j]);
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
                C<R> pj = prev[(int64_t)j * slot_stride + f];
                double cr = coef[2 * j], ci = coef[2 * j + 1];
                vr -= cr * (double)pj.x - ci * (double)pj.y;
                vi -= cr * (double)pj.y + ci * (double)pj.x;
            }
            double inv = coef[2 * nprev];
            v = mk<R>((R)(vr * inv), (R)(vi * inv));
        } else if (mode == EM_SCALE) {
            v = scl<R>(v, (R)shift);
        }
        dst[f] = v;
    }
}

enum RedMode { RM_DOT = 0, RM_DOTX = 1, RM_PMOM = 2, RM_OBS = 3, RM_HCOMB = 4 };

// Per-workgroup partial sums (FP64).  Layout per workgroup: npart doubles.
//  RM_DOT   : sum conj(a) b                      -> 2
//  RM_DOTX  : sum conj(a) x_axis b               -> 2
//  RM_PMOM  : sum |a|^2 (p_j + alpha_j), j<dim   -> dim
//  RM_OBS   : [ |a|^2, |a|^2 x_j, |a|^2 dV/dx_j ] -> 1 + 2*QND_MAXD
//  RM_HCOMB : b <- b_in * fs + V b ; sum conj(a) b  -> 2   (b = H psi, b_in = kinetic part in c)
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
        C<R> va = a[f];
        double ar = va.x, ai = va.y;
        if (mode == RM_DOT || mode == RM_DOTX) {
            C<R> vb = b[f];
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
            double d = ar * ar + ai * ai;
#pragma unroll
            for (int j = 0; j < QND_MAXD; ++j)
                if (j < P.dim) acc[j] += d * (coord_p(P, j, idx[j]) + alpha[j]);
        } else if (mode == RM_OBS) {
            int idx[QND_MAXD];
            flat_index(P, f, idx);
            double rr = 0.0;
#pragma unroll
            for (int j = 0; j < QND_MAXD; ++j) {
                if (j < P.dim) {
                    double x = coord_x(P, j, idx[j]);
                    rr += x * x;
                }
            }
            double dVdu;
            (void)potential(P, rr, &dVdu);
            double d = ar * ar + ai * ai;
            acc[0] += d;
#pragma unroll
            for (int j = 0; j < QND_MAXD; ++j) {
                if (j < P.dim) {
                    double x = coord_x(P, j, idx[j]);
                    acc[1 + j] += d * x;
                    acc[1 + QND_MAXD + j] += d * dVdu * 2.0 * x;
                }
            }
        } else if (mode == RM_HCOMB) {
            int idx[QND_MAXD];
            flat_index(P, f, idx);
            double rr = 0.0;
#pragma unroll
            for (int j = 0; j < QND_MAXD; ++j) {
                if (j < P.dim) {
                    double x = coord_x(P, j, idx[j]);
                    rr += x * x;
                }
            }
            double V = potential(P, rr, nullptr);
            C<R> hk = c[f], hb = b[f];
            double hr = (double)hk.x * fs + V * (double)hb.x;
            double hi = (double)hk.y * fs + V * (double)hb.y;
            b[f] = mk<R>((R)hr, (R)hi);
            acc[0] += ar * hr + ai * hi;
            acc[1] += ar * hi - ai * hr;
        }
    }
    const int nv = (mode == RM_OBS) ? QND_NOBS : (mode == RM_PMOM ? P.dim : 2);
#pragma unroll
    for (int q = 0; q < QND_NOBS; ++q) {
        if (q < nv) {
            double s = block_sum(acc[q], red);
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
        int64_t f = p0 + q;
        int idx[QND_MAXD];
        flat_index(P, f, idx);
        bool inside = true;
        double K = 0.0;
#pragma unroll
        for (int j = 0; j < QND_MAXD; ++j) {
            if (j < P.dim) {
                double p = coord_p(P, j, idx[j]);
                if (fabs(p) > pr_max4[j]) inside = false;
                K += .5 * p * p;
            }
        }
        C<R> v = psi[f];
        dens[q] = inside ? ((double)v.x * (double)v.x + (double)v.y * (double)v.y) : 0.0;
        kin[q] = K;
    }
    __syncthreads();
    const double inv2s2 = 1.0 / (sigma * sigma);
    for (int j = threadIdx.x; j < Ne; j += blockDim.x) {
        double s = 0.0;
        const double Ej = j * de;
        for (int q = 0; q < cnt; ++q) {
            double dk = kin[q] - Ej;
            if (dens[q] != 0.0) s += dens[q] * exp(-.5 * dk * dk * inv2s2) * nf;
        }
        partials[(int64_t)blockIdx.x * Ne + j] = s;
    }
}

}  // namespace qnd
