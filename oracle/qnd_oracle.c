/*
 * qnd_oracle.c -- FP64 CPU restatement of QuantumND's split-operator hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the
 * CPU-baseline leg of the benchmark.  Nothing in the product links it.
 *
 * Each routine restates the reference algorithm (not its code) with the same
 * operation order, FP64 / complex128 throughout, single process (the MPI slab
 * decomposition of the reference with one rank).  The FFT is this file's own
 * (FFTW is not available); it is unnormalised with the FFTW sign convention
 * (forward e^{-2 pi i kn/N}, backward e^{+}), as src/meshgrid.c:166-167 uses.
 *
 * Pinning (what checks each function independently of this file; the tests
 * live in tests/test_oracle.py):
 *   or_fftnd                 numpy.fft (pocketfft) on 1-D/2-D/3-D inputs and
 *                            a brute-force O(N^2) DFT; round trip; Parseval.
 *   or_integrator_coefficients  literals of src/integrators.c:22-56; each half
 *                            sums to 1/2; observed convergence order 2/4/6 on
 *                            the harmonic oscillator.
 *   or_create tables         closed forms of src/meshgrid.c:85-153 (numpy).
 *   or_step_imag             converges to the harmonic-oscillator ground
 *                            energy 1/2 (h^4 for BM4); V=0 free decay.
 *   or_step_length/velocity  free Gaussian packet closed form (V=0, E=0);
 *                            norm conservation without CAP; length/velocity
 *                            gauge equivalence of <x>(t) for a weak field.
 *   or_gram_schmidt          <psi_i|psi_j> = delta_ij to 1e-14; norm-collapse
 *                            error path.
 *   or_energies              dense-matrix Hamiltonian (numpy.linalg.eigh) on
 *                            N <= 64 grids: Rayleigh quotient and <H^2>.
 *   or_observables           displaced Gaussian <x> = x0; boosted Gaussian
 *                            <p> = k0 times the dipole_v scaling of
 *                            src/calculate_observables.c:263; Lz of a real
 *                            state = 0 and of x+iy vortex = 1; dipole_a
 *                            against finite differences of <x>.
 *   or_imag_run              H_1D ground energy (soft-core a^2=2: -0.5 Ha
 *                            design value, loose); orthonormality; step count
 *                            of the FP64 loop (1001 for T=100, dt=0.1).
 *   or_real_run              unitarity; agreement with or_step_* sequence.
 *   or_field_vectors         closed forms; E = -dA/dt by central differences.
 *   or_field_vectors_angle   angle 0 equals or_field_vectors bit for bit; a
 *                            rotated mode equals the rotation matrix applied
 *                            to the unrotated mode's vectors (and |E|, |A|
 *                            unchanged); E = -dA/dt by central differences.
 *   or_final_observables     PMD header counts and order on tiny grids (numpy
 *                            restatement).  PED: the energy sum rule
 *                            sum_j S_j de = (2 pi sigma^2)^((1-d)/2)
 *                            sum_inside |psi_p|^2 (the Gaussian kernel's
 *                            Riemann sum, Poisson error ~ e^{-2 pi^2}), its
 *                            first moment = <K> and its variance = Var(K) +
 *                            sigma^2, with <K>, Var(K) of a boosted Gaussian
 *                            packet from numpy's FFT and from closed forms.
 * The end-to-end composition of the reference (its exact observables.real
 * traces and output files) is not pinned by any reference-held vector: the
 * reference ships none and cannot be built here (FFTW is absent).
 */
#include "qnd_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define OR_MAX_EXP_ARG 700.0   /* src/integrators.h:12 */

struct or_ctx {
    int dim;
    int64_t nr[OR_MAX_DIM];
    double dr[OR_MAX_DIM];
    int64_t N;
    int64_t off[OR_MAX_DIM + 1]; /* concatenated axis offsets (Nc) */
    double *r, *p;               /* concatenated axis grids */
    double *Kin, *Vcap, *Vpot;   /* per point */
    double *dV;                  /* [dim][N] gradient of Vpot (extension) */
    double dNr, dNp, fft_scaling;
    int s;
    double as[2 * OR_MAX_STAGES], as_scaled[2 * OR_MAX_STAGES];
    cplx *in, *out;              /* scratch, like grid->in / grid->out */
    double mask_frac;
    double dt;                   /* DT as given (the imaginary-time loop steps by it) */
    double *mask;                /* per point, extension */
};

/* ------------------------------------------------------------------ */
/* Threads                                                              */
/* ------------------------------------------------------------------ */
/* Pointwise loops in parallel (1, default) or serial (0).  The reference
 * has no parallel pointwise loop: its only threads are FFTW's
 * (src/meshgrid.c:159-161), so a single-rank run is FFT-threaded with
 * serial pointwise passes. */
static int g_pw_parallel = 1;
void or_set_pointwise_parallel(int on) { g_pw_parallel = on != 0; }

void or_set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* ------------------------------------------------------------------ */
/* FFT (own implementation).                                            */
/* Powers of two: iterative radix-2 decimation in time after a          */
/* bit-reversal permutation.  Other lengths (factors 2,3,5,...):        */
/* recursive mixed-radix Cooley-Tukey with an O(p^2) combine.           */
/* ------------------------------------------------------------------ */
typedef struct {
    int64_t n;
    int pow2;
    int64_t *rev;
    cplx *w;        /* w[k] = exp(-2 pi i k / n) */
} fft_plan;

#define OR_PLAN_CACHE 16
static fft_plan g_plans[OR_PLAN_CACHE];
static int g_nplans = 0;

static fft_plan *get_plan(int64_t n)
{
    fft_plan *pl;
    int i;
    for (i = 0; i < g_nplans; i++)
        if (g_plans[i].n == n) return &g_plans[i];
    pl = &g_plans[g_nplans < OR_PLAN_CACHE ? g_nplans++ : OR_PLAN_CACHE - 1];
    free(pl->rev);
    free(pl->w);
    pl->n = n;
    pl->pow2 = (n & (n - 1)) == 0;
    pl->w = (cplx *)malloc((size_t)n * sizeof(cplx));
    for (int64_t k = 0; k < n; k++) {
        long double ang = -2.0L * 3.141592653589793238462643383279502884L *
                          (long double)k / (long double)n;
        pl->w[k] = (double)cosl(ang) + I * (double)sinl(ang);
    }
    pl->rev = NULL;
    if (pl->pow2) {
        int bits = 0;
        while (((int64_t)1 << bits) < n) bits++;
        pl->rev = (int64_t *)malloc((size_t)n * sizeof(int64_t));
        for (int64_t k = 0; k < n; k++) {
            int64_t x = k, y = 0;
            for (int b = 0; b < bits; b++) { y = (y << 1) | (x & 1); x >>= 1; }
            pl->rev[k] = y;
        }
    }
    return pl;
}

static void fft_pow2(const fft_plan *pl, cplx *a, int sign)
{
    int64_t n = pl->n;
    for (int64_t k = 0; k < n; k++) {
        int64_t j = pl->rev[k];
        if (j > k) { cplx t = a[k]; a[k] = a[j]; a[j] = t; }
    }
    for (int64_t len = 2; len <= n; len <<= 1) {
        int64_t half = len >> 1, step = n / len;
        for (int64_t i = 0; i < n; i += len) {
            for (int64_t k = 0; k < half; k++) {
                cplx w = pl->w[k * step];
                if (sign > 0) w = conj(w);
                cplx t = w * a[i + k + half];
                a[i + k + half] = a[i + k] - t;
                a[i + k] = a[i + k] + t;
            }
        }
    }
}

/* out[k] = sum_j in[j*stride] W_n^{jk}, W_n = w[N/n] (sign applied). */
static void fft_rec(cplx *out, const cplx *in, int64_t n, int64_t stride,
                    const cplx *w, int64_t Ntot, int sign)
{
    if (n == 1) { out[0] = in[0]; return; }
    int64_t p = 2;
    while (n % p) p++;
    int64_t m = n / p;
    for (int64_t r = 0; r < p; r++)
        fft_rec(out + r * m, in + r * stride, m, stride * p, w, Ntot, sign);
    cplx *tmp = (cplx *)malloc((size_t)p * sizeof(cplx));
    int64_t wstep = Ntot / n;
    for (int64_t k = 0; k < m; k++) {
        for (int64_t q = 0; q < p; q++) {
            cplx acc = 0;
            for (int64_t r = 0; r < p; r++) {
                int64_t e = (r * (k + q * m)) % n;
                cplx tw = w[e * wstep];
                if (sign > 0) tw = conj(tw);
                acc += tw * out[r * m + k];
            }
            tmp[q] = acc;
        }
        for (int64_t q = 0; q < p; q++) out[q * m + k] = tmp[q];
    }
    free(tmp);
}

static void fft1d(const fft_plan *pl, cplx *a, cplx *scratch, int sign)
{
    if (pl->pow2) { fft_pow2(pl, a, sign); return; }
    fft_rec(scratch, a, pl->n, 1, pl->w, pl->n, sign);
    memcpy(a, scratch, (size_t)pl->n * sizeof(cplx));
}

#define OR_COLBLOCK 8
void or_fftnd(double *data, int dim, const int64_t *nr, int sign)
{
    cplx *a = (cplx *)data;
    int64_t N = 1;
    for (int j = 0; j < dim; j++) N *= nr[j];
    for (int ax = 0; ax < dim; ax++) {
        int64_t n = nr[ax];
        if (n == 1) continue;
        int64_t inner = 1, outer = 1;
        for (int j = ax + 1; j < dim; j++) inner *= nr[j];
        for (int j = 0; j < ax; j++) outer *= nr[j];
        fft_plan *pl;
#ifdef _OPENMP
#pragma omp critical(or_plan)
#endif
        pl = get_plan(n);
        if (inner == 1) {
#ifdef _OPENMP
#pragma omp parallel
#endif
            {
                cplx *scr = (cplx *)malloc((size_t)n * sizeof(cplx));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
                for (int64_t o = 0; o < outer; o++) fft1d(pl, a + o * n, scr, sign);
                free(scr);
            }
        } else {
            int64_t nblk = (inner + OR_COLBLOCK - 1) / OR_COLBLOCK;
#ifdef _OPENMP
#pragma omp parallel
#endif
            {
                cplx *buf = (cplx *)malloc((size_t)n * OR_COLBLOCK * sizeof(cplx));
                cplx *scr = (cplx *)malloc((size_t)n * sizeof(cplx));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
                for (int64_t t = 0; t < outer * nblk; t++) {
                    int64_t o = t / nblk, b0 = (t % nblk) * OR_COLBLOCK;
                    int64_t nb = inner - b0 < OR_COLBLOCK ? inner - b0 : OR_COLBLOCK;
                    cplx *base = a + o * n * inner + b0;
                    for (int64_t i = 0; i < n; i++)
                        for (int64_t b = 0; b < nb; b++) buf[b * n + i] = base[i * inner + b];
                    for (int64_t b = 0; b < nb; b++) fft1d(pl, buf + b * n, scr, sign);
                    for (int64_t i = 0; i < n; i++)
                        for (int64_t b = 0; b < nb; b++) base[i * inner + b] = buf[b * n + i];
                }
                free(buf);
                free(scr);
            }
        }
    }
    (void)N;
}

/* ------------------------------------------------------------------ */
/* Integrator (src/integrators.c:13-71)                                 */
/* ------------------------------------------------------------------ */
int or_integrator_coefficients(const char *name, double dt, double *as,
                               double *as_scaled)
{
    int s, j;
    for (j = 0; j < 2 * OR_MAX_STAGES; j++) { as[j] = 0.0; as_scaled[j] = 0.0; }
    if (strcmp(name, "Verlet") == 0) {
        s = 1; as[0] = .5;
    } else if (strcmp(name, "BM4") == 0) {
        static const double c[6] = {.0792036964311957, .1303114101821663,
                                    .2228614958676077, -.3667132690474257,
                                    .3246481886897062, .1096884778767498};
        s = 6; memcpy(as, c, sizeof c);
    } else if (strcmp(name, "BM6") == 0) {
        static const double c[10] = {.050262764400392, .098553683500650,
                                     .314960616927694, -.447346482695478,
                                     .492426372489876, -.425118767797691,
                                     .237063913978122, .195602488600053,
                                     .346358189850727, -.362762779254345};
        s = 10; memcpy(as, c, sizeof c);
    } else if (strcmp(name, "RKN4") == 0) {
        static const double c[6] = {.082984406417405, .162314550766866,
                                    .233995250731502, .370877414979578,
                                    -.409933719901926, .059762097006575};
        s = 6; memcpy(as, c, sizeof c);
    } else {
        return -1;
    }
    /* the second half mirrors the first */
    for (j = 0; j < s; j++) as[s + j] = as[s - 1 - j];
    for (j = 0; j < 2 * s; j++) as_scaled[j] = as[j] * dt;
    return s;
}

/* ------------------------------------------------------------------ */
/* Grid and tables (src/meshgrid.c:17-179)                              */
/* ------------------------------------------------------------------ */
enum { POT_SOFT = 0, POT_GAUSS = 1, POT_ZERO = 2, POT_HARM = 3 };

static int potential_params(const char *name, int *kind, double *b, double *c2)
{
    /* V = -(1 + b e^{-r^2}) / sqrt(r^2 + c2)  (kind SOFT)
     * V = -2.7459 e^{-r^2/(2*1.96)}           (kind GAUSS) */
    *kind = POT_SOFT; *b = 0.0;
    if (strcmp(name, "H_1D") == 0) *c2 = 2.;
    else if (strcmp(name, "H_3D") == 0) *c2 = 1e-4;
    else if (strcmp(name, "H_2D") == 0) *c2 = .7396;
    else if (strcmp(name, "Ne_2D") == 0) { *b = 9.; *c2 = 2.88172; }
    else if (strcmp(name, "Ar_2D") == 0) { *b = 17.; *c2 = .91205; }
    else if (strcmp(name, "Kr_2D") == 0) { *b = 35.; *c2 = .7875; }
    else if (strcmp(name, "Xe_2D") == 0) { *b = 53.; *c2 = .52592; }
    else if (strcmp(name, "Ne_short_2D") == 0) { *kind = POT_GAUSS; *c2 = 0; }
    /* pinning-only potentials (closed-form tests of the steps, not in the
     * reference): V = 0 and V = r^2/2 */
    else if (strcmp(name, "zero_test") == 0) { *kind = POT_ZERO; *c2 = 0; }
    else if (strcmp(name, "harmonic_test") == 0) { *kind = POT_HARM; *c2 = 0; }
    else return -1;
    return 0;
}

or_ctx *or_create(int dim, const int64_t *nr, const double *dr,
                  const char *potential, double softcore, const double *cap,
                  const char *integrator, double dt, char *err, int errlen)
{
    int kind;
    double b, c2;
    if (dim < 1 || dim > OR_MAX_DIM) {
        if (err) snprintf(err, errlen, "bad dimension %d", dim);
        return NULL;
    }
    if (potential_params(potential, &kind, &b, &c2) != 0) {
        if (err) snprintf(err, errlen, "Unknown potential: %s", potential);
        return NULL;
    }
    if (softcore >= 0) c2 = softcore * softcore;
    or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
    c->dim = dim;
    c->N = 1;
    c->off[0] = 0;
    for (int j = 0; j < dim; j++) {
        c->nr[j] = nr[j];
        c->dr[j] = dr[j];
        c->N *= nr[j];
        c->off[j + 1] = c->off[j] + nr[j];
    }
    c->s = or_integrator_coefficients(integrator, dt, c->as, c->as_scaled);
    c->dt = dt;
    if (c->s < 0) {
        if (err) snprintf(err, errlen, "Unknown integrator: %s", integrator);
        free(c);
        return NULL;
    }
    /* volumes (src/meshgrid.c:23-28) */
    c->dNr = 1.;
    c->dNp = 1.;
    for (int j = 0; j < dim; j++) {
        c->dNr *= dr[j];
        c->dNp *= 2. * M_PI / (double)(dr[j] * nr[j]);
    }
    /* axis grids, p in FFT order (src/meshgrid.c:76-89) */
    c->r = (double *)malloc((size_t)c->off[dim] * sizeof(double));
    c->p = (double *)malloc((size_t)c->off[dim] * sizeof(double));
    for (int j = 0; j < dim; j++) {
        double L = dr[j] * nr[j];
        double dp = 2. * M_PI / L;
        for (int64_t i = 0; i < nr[j]; i++) {
            c->r[c->off[j] + i] = -.5 * L + i * dr[j];
            c->p[c->off[j] + i] = (i < nr[j] / 2) ? i * dp : (i - nr[j]) * dp;
        }
    }
    int64_t N = c->N;
    c->Kin = (double *)malloc((size_t)N * sizeof(double));
    c->Vcap = (double *)malloc((size_t)N * sizeof(double));
    c->Vpot = (double *)malloc((size_t)N * sizeof(double));
    c->dV = (double *)malloc((size_t)N * dim * sizeof(double));
    c->mask = NULL;
    c->mask_frac = 0;
    double capv[2] = {0, 0};
    if (cap) { capv[0] = cap[0]; capv[1] = cap[1]; }
    /* per-point tables (src/meshgrid.c:123-154) */
    for (int64_t i = 0; i < N; i++) {
        int64_t idx[OR_MAX_DIM], rem = i;
        for (int j = dim - 1; j >= 0; j--) { idx[j] = rem % nr[j]; rem /= nr[j]; }
        double K = 0., Vc = 0., rr = 0.;
        double xs[OR_MAX_DIM];
        for (int j = 0; j < dim; j++) {
            double x = c->r[c->off[j] + idx[j]];
            double pj = c->p[c->off[j] + idx[j]];
            xs[j] = x;
            K += .5 * pj * pj;
            rr += x * x;
            double Lc = .5 * nr[j] * dr[j] - capv[1];
            if (fabs(x) >= Lc) Vc += capv[0] * (fabs(x) - Lc) * (fabs(x) - Lc);
        }
        double rad = sqrt(rr);
        double u = rad * rad;
        double V, dVdu;
        if (kind == POT_SOFT) {
            double den = sqrt(u + c2);
            double g = (b != 0.) ? b * exp(-u) : 0.;
            V = -(1. + g) / den;
            dVdu = g / den + .5 * (1. + g) / (den * (u + c2));
        } else if (kind == POT_GAUSS) {
            double g = 2.7459 * exp(-.5 * rad * rad / 1.96);
            V = -g;
            dVdu = g * .5 / 1.96;
        } else if (kind == POT_ZERO) {
            V = 0.;
            dVdu = 0.;
        } else {
            V = .5 * u;
            dVdu = .5;
        }
        c->Kin[i] = K;
        c->Vcap[i] = Vc;
        c->Vpot[i] = V;
        for (int j = 0; j < dim; j++) c->dV[(int64_t)j * N + i] = dVdu * 2. * xs[j];
    }
    /* fft scaling (src/meshgrid.c:162-163) */
    c->fft_scaling = 1.;
    for (int j = 0; j < dim; j++) c->fft_scaling /= (double)nr[j];
    c->in = (cplx *)calloc((size_t)N, sizeof(cplx));
    c->out = (cplx *)calloc((size_t)N, sizeof(cplx));
    return c;
}

void or_destroy(or_ctx *c)
{
    if (!c) return;
    free(c->r); free(c->p); free(c->Kin); free(c->Vcap); free(c->Vpot);
    free(c->dV); free(c->in); free(c->out); free(c->mask);
    free(c);
}

int64_t or_npoints(const or_ctx *c) { return c->N; }
int or_stages(const or_ctx *c) { return c->s; }
void or_coefficients(const or_ctx *c, double *as, double *as_scaled)
{
    memcpy(as, c->as, sizeof c->as);
    memcpy(as_scaled, c->as_scaled, sizeof c->as_scaled);
}
void or_tables(const or_ctx *c, double *r, double *p, double *Kin, double *Vcap,
               double *Vpot)
{
    if (r) memcpy(r, c->r, (size_t)c->off[c->dim] * sizeof(double));
    if (p) memcpy(p, c->p, (size_t)c->off[c->dim] * sizeof(double));
    if (Kin) memcpy(Kin, c->Kin, (size_t)c->N * sizeof(double));
    if (Vcap) memcpy(Vcap, c->Vcap, (size_t)c->N * sizeof(double));
    if (Vpot) memcpy(Vpot, c->Vpot, (size_t)c->N * sizeof(double));
}
void or_scalings(const or_ctx *c, double *dNr, double *dNp, double *fs)
{
    *dNr = c->dNr; *dNp = c->dNp; *fs = c->fft_scaling;
}

static inline double axis_r(const or_ctx *c, int j, int64_t i)
{
    int64_t rem = i;
    for (int k = c->dim - 1; k > j; k--) rem /= c->nr[k];
    return c->r[c->off[j] + rem % c->nr[j]];
}
static inline double axis_p(const or_ctx *c, int j, int64_t i)
{
    int64_t rem = i;
    for (int k = c->dim - 1; k > j; k--) rem /= c->nr[k];
    return c->p[c->off[j] + rem % c->nr[j]];
}

static void fwd(or_ctx *c) { memcpy(c->out, c->in, (size_t)c->N * sizeof(cplx)); or_fftnd((double *)c->out, c->dim, c->nr, -1); }
static void bwd(or_ctx *c) { memcpy(c->out, c->in, (size_t)c->N * sizeof(cplx)); or_fftnd((double *)c->out, c->dim, c->nr, +1); }

static inline double clampexp(double a) { return a > OR_MAX_EXP_ARG ? OR_MAX_EXP_ARG : a; }

/* ------------------------------------------------------------------ */
/* Steps (src/integrators.c:135-232)                                    */
/* ------------------------------------------------------------------ */
void or_step_imag(or_ctx *c, double *psi_d)
{
    cplx *psi = (cplx *)psi_d;
    int64_t N = c->N;
    for (int ia = 0; ia < c->s; ia++) {
        double a0 = c->as_scaled[2 * ia], a1 = c->as_scaled[2 * ia + 1];
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (g_pw_parallel)
#endif
        for (int64_t i = 0; i < N; i++)
            c->in[i] = exp(clampexp(-a0 * c->Vpot[i])) * psi[i];
        fwd(c);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (g_pw_parallel)
#endif
        for (int64_t i = 0; i < N; i++)
            c->in[i] = exp(clampexp(-(a0 + a1) * c->Kin[i])) * c->out[i];
        bwd(c);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (g_pw_parallel)
#endif
        for (int64_t i = 0; i < N; i++)
            psi[i] = exp(clampexp(-a1 * c->Vpot[i])) * c->out[i] * c->fft_scaling;
    }
}

void or_step_length(or_ctx *c, double *psi_d, const double *Et)
{
    cplx *psi = (cplx *)psi_d;
    int64_t N = c->N;
    int dim = c->dim;
    for (int ia = 0; ia < c->s; ia++) {
        double a0 = c->as_scaled[2 * ia], a1 = c->as_scaled[2 * ia + 1];
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (g_pw_parallel)
#endif
        for (int64_t i = 0; i < N; i++)
            c->in[i] = cexp(-I * a0 * c->Kin[i]) * psi[i];
        bwd(c);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (g_pw_parallel)
#endif
        for (int64_t i = 0; i < N; i++) {
            double W = 0.;
            for (int j = 0; j < dim; j++) W += axis_r(c, j, i) * Et[ia * dim + j];
            double arg = clampexp(-(a0 + a1) * c->Vcap[i]);
            c->in[i] = cexp(arg - I * (a0 + a1) * (c->Vpot[i] + W)) * c->out[i];
        }
        fwd(c);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (g_pw_parallel)
#endif
        for (int64_t i = 0; i < N; i++)
            psi[i] = cexp(-I * a1 * c->Kin[i]) * c->out[i] * c->fft_scaling;
    }
}

void or_step_velocity(or_ctx *c, double *psi_d, const double *At)
{
    cplx *psi = (cplx *)psi_d;
    int64_t N = c->N;
    int dim = c->dim;
    for (int ia = 0; ia < c->s; ia++) {
        double a0 = c->as_scaled[2 * ia], a1 = c->as_scaled[2 * ia + 1];
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (g_pw_parallel)
#endif
        for (int64_t i = 0; i < N; i++) {
            double arg = clampexp(-a0 * c->Vcap[i]);
            c->in[i] = cexp(arg - I * a0 * c->Vpot[i]) * psi[i];
        }
        fwd(c);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (g_pw_parallel)
#endif
        for (int64_t i = 0; i < N; i++) {
            double W = 0.;
            for (int j = 0; j < dim; j++) {
                double Aj = At[ia * dim + j];
                W += (axis_p(c, j, i) + .5 * Aj) * Aj;
            }
            c->in[i] = cexp(-I * (a0 + a1) * (c->Kin[i] + W)) * c->out[i];
        }
        bwd(c);
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (g_pw_parallel)
#endif
        for (int64_t i = 0; i < N; i++) {
            double arg = clampexp(-a1 * c->Vcap[i]);
            psi[i] = cexp(arg - I * a1 * c->Vpot[i]) * c->out[i] * c->fft_scaling;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Gram-Schmidt (src/QuantumND_imag.c:265-311)                          */
/* ------------------------------------------------------------------ */
int or_gram_schmidt(or_ctx *c, int Nst, double *states)
{
    int64_t N = c->N;
    cplx *all = (cplx *)states;
    cplx *target = all + (int64_t)Nst * N;
    cplx *v = (cplx *)malloc((size_t)N * sizeof(cplx));
    memcpy(v, target, (size_t)N * sizeof(cplx));
    for (int j = 0; j < Nst; j++) {
        const cplx *pj = all + (int64_t)j * N;
        cplx sc = 0.;
        for (int64_t i = 0; i < N; i++) sc += conj(pj[i]) * v[i] * c->dNr;
        for (int64_t i = 0; i < N; i++) v[i] -= sc * pj[i];
    }
    double nrm = 0.;
    for (int64_t i = 0; i < N; i++) nrm += cabs(v[i]) * cabs(v[i]) * c->dNr;
    if (nrm < 1.e-12) { free(v); return 1; }
    double inv = 1. / sqrt(nrm);
    for (int64_t i = 0; i < N; i++) target[i] = v[i] * inv;
    free(v);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Observables (src/calculate_observables.c:89-266)                     */
/* ------------------------------------------------------------------ */
static void local_energies(or_ctx *c, const cplx *psi, int M, cplx *E)
{
    int64_t N = c->N;
    cplx *H = (cplx *)malloc((size_t)N * sizeof(cplx));
    memcpy(H, psi, (size_t)N * sizeof(cplx));
    for (int m = 0; m < M; m++) {
        E[m] = 0.;
        memcpy(c->in, H, (size_t)N * sizeof(cplx));
        fwd(c);
        for (int64_t i = 0; i < N; i++) c->in[i] = c->Kin[i] * c->out[i];
        bwd(c);
        for (int64_t i = 0; i < N; i++) {
            H[i] = c->out[i] * c->fft_scaling + c->Vpot[i] * H[i];
            E[m] += conj(psi[i]) * H[i] * c->dNr;
        }
    }
    free(H);
}

void or_energies(or_ctx *c, const double *psi, int M, double *E)
{
    cplx *Ec = (cplx *)malloc((size_t)M * sizeof(cplx));
    local_energies(c, (const cplx *)psi, M, Ec);
    for (int m = 0; m < M; m++) E[m] = creal(Ec[m]);
    free(Ec);
}

int or_obs_width(int code, int dim)
{
    switch (code) {
    case OR_OBS_LZ: case OR_OBS_ENERGY: case OR_OBS_NORM: return 1;
    case OR_OBS_DIPOLE_R: case OR_OBS_DIPOLE_V: case OR_OBS_DIPOLE_A: return dim;
    default: return 0;
    }
}

static cplx local_Lz(or_ctx *c, const cplx *psi, const double *alpha)
{
    int64_t N = c->N;
    cplx Lz = 0.;
    cplx *pp = (cplx *)malloc((size_t)N * sizeof(cplx));
    memcpy(c->in, psi, (size_t)N * sizeof(cplx));
    fwd(c);
    memcpy(pp, c->out, (size_t)N * sizeof(cplx));
    double Ax = alpha[0], Ay = alpha[1];
    for (int64_t i = 0; i < N; i++) c->in[i] = (axis_p(c, 1, i) + Ay) * pp[i];
    bwd(c);
    for (int64_t i = 0; i < N; i++) Lz += conj(psi[i]) * axis_r(c, 0, i) * c->out[i];
    for (int64_t i = 0; i < N; i++) c->in[i] = (axis_p(c, 0, i) + Ax) * pp[i];
    bwd(c);
    for (int64_t i = 0; i < N; i++) Lz -= conj(psi[i]) * axis_r(c, 1, i) * c->out[i];
    Lz *= c->fft_scaling * c->dNr;
    free(pp);
    return Lz;
}

void or_observables(or_ctx *c, const double *psi_d, const int *codes, int ncodes,
                    const double *alpha, const double *Efield, double *out)
{
    const cplx *psi = (const cplx *)psi_d;
    int64_t N = c->N;
    int dim = c->dim;
    int off = 0;
    for (int k = 0; k < ncodes; k++) {
        int code = codes[k];
        if (code == OR_OBS_LZ) {
            out[off] = (dim <= 1) ? 0. : creal(local_Lz(c, psi, alpha));
        } else if (code == OR_OBS_ENERGY) {
            cplx e;
            local_energies(c, psi, 1, &e);
            out[off] = creal(e);
        } else if (code == OR_OBS_DIPOLE_R) {
            for (int j = 0; j < dim; j++) {
                cplx d = 0.;
                for (int64_t i = 0; i < N; i++)
                    d += conj(psi[i]) * axis_r(c, j, i) * psi[i] * c->dNr;
                out[off + j] = creal(d);
            }
        } else if (code == OR_OBS_DIPOLE_V) {
            memcpy(c->in, psi, (size_t)N * sizeof(cplx));
            fwd(c);
            for (int j = 0; j < dim; j++) {
                cplx d = 0.;
                for (int64_t i = 0; i < N; i++)
                    d += conj(c->out[i]) * (axis_p(c, j, i) + alpha[j]) * c->out[i] *
                         c->dNp * c->fft_scaling;
                out[off + j] = creal(d);
            }
        } else if (code == OR_OBS_NORM) {
            double s = 0.;
            for (int64_t i = 0; i < N; i++) s += cabs(psi[i]) * cabs(psi[i]) * c->dNr;
            out[off] = s;
        } else if (code == OR_OBS_DIPOLE_A) {
            double nrm = 0.;
            for (int64_t i = 0; i < N; i++) nrm += cabs(psi[i]) * cabs(psi[i]) * c->dNr;
            for (int j = 0; j < dim; j++) {
                double s = 0.;
                for (int64_t i = 0; i < N; i++)
                    s += cabs(psi[i]) * cabs(psi[i]) * c->dV[(int64_t)j * N + i] * c->dNr;
                out[off + j] = -s - (Efield ? Efield[j] : 0.) * nrm;
            }
        }
        off += or_obs_width(code, dim);
    }
}

/* ------------------------------------------------------------------ */
/* Mask extension: per-axis cos^(1/8) ramp over the outer `frac` of the */
/* half-width L_j = n_j dr_j / 2.                                       */
/* ------------------------------------------------------------------ */
void or_set_mask(or_ctx *c, double frac)
{
    free(c->mask);
    c->mask = NULL;
    c->mask_frac = frac;
    if (frac <= 0) return;
    int64_t N = c->N;
    c->mask = (double *)malloc((size_t)N * sizeof(double));
    for (int64_t i = 0; i < N; i++) {
        double m = 1.;
        for (int j = 0; j < c->dim; j++) {
            double L = .5 * c->nr[j] * c->dr[j];
            double x0 = (1. - frac) * L;
            double ax = fabs(axis_r(c, j, i));
            if (ax > x0) m *= pow(cos(.5 * M_PI * (ax - x0) / (L - x0)), .125);
        }
        c->mask[i] = m;
    }
}

void or_apply_mask(or_ctx *c, double *psi_d)
{
    if (!c->mask) return;
    cplx *psi = (cplx *)psi_d;
    for (int64_t i = 0; i < c->N; i++) psi[i] *= c->mask[i];
}

/* ------------------------------------------------------------------ */
/* Initial guesses                                                      */
/* ------------------------------------------------------------------ */
static uint64_t splitmix64(uint64_t *state)
{
    uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void or_gaussian(uint64_t seed, int64_t n, double *out)
{
    uint64_t st = seed;
    for (int64_t i = 0; i < n; i += 2) {
        double u1 = ((double)(splitmix64(&st) >> 11) + 1.0) * 0x1.0p-53;
        double u2 = ((double)(splitmix64(&st) >> 11) + 1.0) * 0x1.0p-53;
        double rad = sqrt(-2.0 * log(u1));
        out[i] = rad * cos(2.0 * M_PI * u2);
        if (i + 1 < n) out[i + 1] = rad * sin(2.0 * M_PI * u2);
    }
}

void or_initial_guess(or_ctx *c, int n_states, int guess_mode, uint64_t seed,
                      double *psi_out)
{
    int64_t N = c->N;
    double *g = NULL;
    if (guess_mode == 1) {
        g = (double *)malloc((size_t)N * 2 * n_states * sizeof(double));
        or_gaussian(seed, N * 2 * n_states, g);
    }
    for (int st = 0; st < n_states; st++) {
        cplx *psi = (cplx *)psi_out + (int64_t)st * N;
        for (int64_t i = 0; i < N; i++) {
            double rr = 0.;
            for (int k = 0; k < c->dim; k++) {
                double x = axis_r(c, k, i);
                rr += x * x;
            }
            if (guess_mode == 0) {
                /* src/QuantumND_imag.c:174-179 */
                double rad = sqrt(rr);
                psi[i] = (rad + (double)(i % 3)) * exp(-rad) + I * 0.;
            } else {
                const double *gi = g + ((int64_t)st * N + i) * 2;
                psi[i] = (gi[0] + I * gi[1]) * exp(-rr / 20.);
            }
        }
    }
    free(g);
}

int64_t or_imag_step_count(double T, double dt)
{
    int64_t n = 0;
    for (double t = 0.; t < T; t += dt) n++;
    return n;
}

/* ------------------------------------------------------------------ */
/* Imaginary-time driver (src/QuantumND_imag.c:169-240)                 */
/* ------------------------------------------------------------------ */
int or_imag_run(or_ctx *c, int n_states, double T, int guess_mode,
                uint64_t seed, double *psi_out, double *E_out, double *Lz_out)
{
    int64_t N = c->N;
    double dt = c->dt;
    double alpha[OR_MAX_DIM] = {0};
    or_initial_guess(c, n_states, guess_mode, seed, psi_out);
    for (int st = 0; st < n_states; st++) {
        double *psi = psi_out + (int64_t)st * N * 2;
        for (double t = 0.; t < T; t += dt) {
            or_step_imag(c, psi);
            if (or_gram_schmidt(c, st, psi_out) != 0) return 1;
        }
        double E2[2];
        or_energies(c, psi, 2, E2);
        E_out[2 * st] = E2[0];
        E_out[2 * st + 1] = E2[1];
        int code = OR_OBS_LZ;
        or_observables(c, psi, &code, 1, alpha, NULL, &Lz_out[st]);
    }
    /* global phase at the first strict maximum of |psi| */
    for (int st = 0; st < n_states; st++) {
        cplx *psi = (cplx *)psi_out + (int64_t)st * N;
        double rho = 0., ph = 0.;
        for (int64_t i = 0; i < N; i++) {
            double m = cabs(psi[i]);
            if (m > rho) { rho = m; ph = carg(psi[i]); }
        }
        cplx f = cexp(-I * ph);
        for (int64_t i = 0; i < N; i++) psi[i] *= f;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Real-time driver (src/QuantumND_real.c:166-228)                      */
/* ------------------------------------------------------------------ */
int or_real_run(or_ctx *c, int gauge, double *psi_d, int64_t ntime,
                const double *Et, const double *At, const double *Eobs,
                const double *Aobs, const int *codes, int ncodes, double *out)
{
    cplx *psi = (cplx *)psi_d;
    int64_t N = c->N;
    int dim = c->dim, s = c->s;
    int width = 0;
    for (int k = 0; k < ncodes; k++) width += or_obs_width(codes[k], dim);
    double zero[OR_MAX_DIM] = {0};
    double sq = sqrt(c->fft_scaling);
    if (gauge == 1) {
        for (int64_t it = 0; it < ntime; it++) {
            if (ncodes > 0)
                or_observables(c, psi_d, codes, ncodes, Aobs + it * dim,
                               Eobs ? Eobs + it * dim : NULL, out + it * width);
            or_step_velocity(c, psi_d, At + it * s * dim);
            or_apply_mask(c, psi_d);
        }
        return 0;
    }
    /* length gauge: psi held in p-space with unitary scaling */
    memcpy(c->in, psi, (size_t)N * sizeof(cplx));
    fwd(c);
    for (int64_t i = 0; i < N; i++) psi[i] = c->out[i] * sq;
    for (int64_t it = 0; it < ntime; it++) {
        memcpy(c->in, psi, (size_t)N * sizeof(cplx));
        bwd(c);
        for (int64_t i = 0; i < N; i++) psi[i] = c->out[i] * sq;
        if (it > 0) or_apply_mask(c, psi_d);     /* mask after the previous step */
        if (ncodes > 0)
            or_observables(c, psi_d, codes, ncodes, zero,
                           Eobs ? Eobs + it * dim : NULL, out + it * width);
        memcpy(c->in, psi, (size_t)N * sizeof(cplx));
        fwd(c);
        for (int64_t i = 0; i < N; i++) psi[i] = c->out[i] * sq;
        or_step_length(c, psi_d, Et + it * s * dim);
    }
    memcpy(c->in, psi, (size_t)N * sizeof(cplx));
    bwd(c);
    for (int64_t i = 0; i < N; i++) psi[i] = c->out[i] * sq;
    if (ntime > 0) or_apply_mask(c, psi_d);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Field (src/parameters.c:899-949, src/integrators.c:74-129)           */
/* ------------------------------------------------------------------ */
int or_field_vectors(int dim, int nmodes, const double *freq, const double *amp,
                     const double *ellip, const double *duration,
                     const double *cep, const double *delay, const int *env,
                     double t, double *A, double *E)
{
    return or_field_vectors_angle(dim, nmodes, freq, amp, ellip, duration, cep, delay, env, NULL, t, A, E);
}

int or_field_vectors_angle(int dim, int nmodes, const double *freq, const double *amp,
                           const double *ellip, const double *duration,
                           const double *cep, const double *delay, const int *env,
                           const double *angle, double t, double *A, double *E)
{
    double f = 0., df = 0.;
    for (int j = 0; j < dim; j++) { E[j] = 0.; A[j] = 0.; }
    for (int j = 0; j < nmodes; j++) {
        double w = freq[j], F0 = amp[j], xi = ellip[j], tau = duration[j];
        double ph = cep[j], ts = t - delay[j];
        if (env[j] == 0) {
            f = 1.; df = 0.;
        } else if (env[j] == 1) {
            double tr = 2. * 2. * M_PI / w;
            f = pow(sin(.5 * M_PI * ts / tr), 4) * (ts >= 0. && ts <= tr) +
                (ts > tr && ts < (tau - tr)) +
                pow(sin(.5 * M_PI * (ts - tau) / tr), 4) * (ts >= tau - tr && ts <= tau);
            df = M_PI / tr *
                 (pow(sin(.5 * M_PI * ts / tr), 2) * sin(M_PI * ts / tr) * (ts >= 0. && ts <= tr) +
                  pow(sin(.5 * M_PI * (ts - tau) / tr), 2) * sin(M_PI * (ts - tau) / tr) *
                      (ts >= tau - tr && ts <= tau));
        } else if (env[j] == 2) {
            f = pow(sin(M_PI * ts / tau), 2) * (ts >= 0. && ts <= tau);
            df = M_PI * sin(2. * M_PI * ts / tau) / tau * (ts >= 0. && ts <= tau);
        } else if (env[j] == 3) {
            f = pow(sin(M_PI * ts / tau), 4) * (ts >= 0. && ts <= tau);
            df = 4. * M_PI * cos(M_PI * ts / tau) * pow(sin(M_PI * ts / tau), 3) / tau *
                 (ts >= 0. && ts <= tau);
        } else {
            return 1;
        }
        if (!angle) {  /* the reference's accumulation, term by term */
            E[0] += -F0 * (df * cos(w * ts + ph) - f * w * sin(w * ts + ph)) / w;
            if (dim >= 2) E[1] += -F0 * xi * (df * sin(w * ts + ph) + f * w * cos(w * ts + ph)) / w;
            A[0] += F0 * f * cos(w * ts + ph) / w;
            if (dim >= 2) A[1] += F0 * xi * f * sin(w * ts + ph) / w;
        } else {
            /* extension: the mode's (x, y) vectors rotated by angle[j] in the
             * xy plane (major axis at angle[j] from x) */
            const double ex = -F0 * (df * cos(w * ts + ph) - f * w * sin(w * ts + ph)) / w;
            const double ey = -F0 * xi * (df * sin(w * ts + ph) + f * w * cos(w * ts + ph)) / w;
            const double ax = F0 * f * cos(w * ts + ph) / w;
            const double ay = F0 * xi * f * sin(w * ts + ph) / w;
            const double ca = cos(angle[j]), sa = sin(angle[j]);
            E[0] += ca * ex - sa * ey;
            A[0] += ca * ax - sa * ay;
            if (dim >= 2) {
                E[1] += sa * ex + ca * ey;
                A[1] += sa * ax + ca * ay;
            }
        }
    }
    return 0;
}

int or_field_tables(int dim, int nmodes, const double *freq, const double *amp,
                    const double *ellip, const double *duration,
                    const double *cep, const double *delay, const int *env,
                    const double *as_scaled, int s, const double *time,
                    int64_t ntime, double *A, double *E)
{
    return or_field_tables_angle(dim, nmodes, freq, amp, ellip, duration, cep, delay, env, NULL, as_scaled, s,
                                 time, ntime, A, E);
}

int or_field_tables_angle(int dim, int nmodes, const double *freq, const double *amp,
                          const double *ellip, const double *duration,
                          const double *cep, const double *delay, const int *env,
                          const double *angle, const double *as_scaled, int s,
                          const double *time, int64_t ntime, double *A, double *E)
{
    for (int64_t i = 0; i < ntime; i++) {
        double t = time[i] + as_scaled[0];
        int64_t ind = i * dim * s;
        for (int j = 0; j < s; j++) {
            if (or_field_vectors_angle(dim, nmodes, freq, amp, ellip, duration, cep, delay,
                                       env, angle, t, &A[ind + j * dim], &E[ind + j * dim]))
                return 1;
            if (j < s - 1) t += as_scaled[2 * j + 1] + as_scaled[2 * j + 2];
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Final observables (src/io_functions.c:136-287)                       */
/* ------------------------------------------------------------------ */
int64_t or_final_observables(or_ctx *c, double *psi_d, const double *R_mask,
                             const double *pr_max, int Ne, double de, double sigma,
                             double *ped, double *pmd, int64_t *npr)
{
    cplx *psi = (cplx *)psi_d;
    int64_t N = c->N;
    int dim = c->dim;
    for (int64_t i = 0; i < N; i++) {
        double rr = 0.;
        for (int j = 0; j < dim; j++) {
            double x = axis_r(c, j, i);
            rr += x * x;
        }
        double rad = sqrt(rr);
        double fm = 0.;
        if (rad >= R_mask[0] && rad <= R_mask[1])
            fm = pow(sin(.5 * M_PI * (rad - R_mask[0]) / (R_mask[1] - R_mask[0])), 2.);
        else if (rad > R_mask[1])
            fm = 1.;
        c->in[i] = fm * psi[i];
    }
    fwd(c);
    double sq = sqrt(c->fft_scaling);
    for (int64_t i = 0; i < N; i++) psi[i] = c->out[i] * sq;
    if (ped) {
        double nf = 1. / pow(2. * M_PI * sigma * sigma, .5 * dim);
        for (int j = 0; j < Ne; j++) ped[j] = 0.;
        for (int64_t i = 0; i < N; i++) {
            int j;
            for (j = 0; j < dim; j++)
                if (fabs(axis_p(c, j, i)) > pr_max[j]) break;
            if (j == dim)
                for (j = 0; j < Ne; j++)
                    ped[j] += cabs(psi[i]) * cabs(psi[i]) *
                              exp(-.5 * (c->Kin[i] - j * de) * (c->Kin[i] - j * de) /
                                  (sigma * sigma)) * nf;
        }
    }
    if (npr)
        for (int j = 0; j < dim; j++) {
            npr[j] = 0;
            for (int64_t i = 0; i < c->nr[j]; i++)
                if (fabs(c->p[c->off[j] + i]) <= pr_max[j]) npr[j]++;
        }
    int64_t cnt = 0;
    for (int64_t i = 0; i < N; i++) {
        int j;
        for (j = 0; j < dim; j++)
            if (fabs(axis_p(c, j, i)) > pr_max[j]) break;
        if (j == dim) {
            if (pmd) {
                pmd[2 * cnt] = cabs(psi[i]) * cabs(psi[i]);
                pmd[2 * cnt + 1] = carg(psi[i]);
            }
            cnt++;
        }
    }
    return cnt;
}
