/*
 * qnd_oracle.h -- FP64 CPU restatement of QuantumND's split-operator hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark call it; the
 * product (libqnd and the QuantumND_* executables) never links or calls it.
 *
 * Every function restates the reference algorithm (complex128, FP64
 * throughout, same operation order) from its own code.  Citations are to the
 * reference sources (src/...).  Pinning of each function against independent
 * checks is listed in the header of qnd_oracle.c.
 *
 * Complex arrays cross this interface as interleaved doubles (re, im) in
 * row-major order, axis 0 slowest, exactly the reference layout
 * (src/meshgrid.c:205-211).
 */
#ifndef QND_ORACLE_H
#define QND_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_DIM 4
#define OR_MAX_STAGES 10

typedef struct or_ctx or_ctx;

/* Potential selection: name as in src/meshgrid.c:140-153.  softcore < 0 keeps
 * the reference's constant; softcore >= 0 replaces the constant added to r^2
 * by softcore^2 (extension for the "soft-core a" configurations). */
or_ctx *or_create(int dim, const int64_t *nr, const double *dr,
                  const char *potential, double softcore,
                  const double *cap /* [2] */,
                  const char *integrator, double dt, char *err, int errlen);
void or_destroy(or_ctx *c);
void or_set_threads(int nthreads);         /* 0 = OpenMP default */
/* 1 (default): pointwise loops in parallel; 0: serial, the FFT alone
 * threaded (the reference's single-rank threading, src/meshgrid.c:159-161). */
void or_set_pointwise_parallel(int on);

/* Accessors (for tests). */
int64_t or_npoints(const or_ctx *c);
int or_stages(const or_ctx *c);
void or_coefficients(const or_ctx *c, double *as, double *as_scaled); /* 2*MAX */
void or_tables(const or_ctx *c, double *r, double *p, double *Kin,
               double *Vcap, double *Vpot);   /* r,p: concatenated axes */
void or_scalings(const or_ctx *c, double *dNr, double *dNp, double *fft_scaling);

/* Integrator coefficients only (src/integrators.c:13-71).  Returns s, or -1
 * for an unknown name. */
int or_integrator_coefficients(const char *name, double dt, double *as,
                               double *as_scaled);

/* Unnormalised N-D DFT, sign -1 (forward) or +1 (backward), in place. */
void or_fftnd(double *a, int dim, const int64_t *nr, int sign);

/* One split-operator step (src/integrators.c:135-232). */
void or_step_imag(or_ctx *c, double *psi);
void or_step_length(or_ctx *c, double *psi_p, const double *Et /* [s][dim] */);
void or_step_velocity(or_ctx *c, double *psi, const double *At /* [s][dim] */);

/* Modified Gram-Schmidt of state N against states 0..N-1
 * (src/QuantumND_imag.c:265-311).  states = (N+1) consecutive arrays.
 * Returns 0, or 1 when the norm collapses below 1e-12. */
int or_gram_schmidt(or_ctx *c, int N, double *states);

/* Energy moments <H^m>, m=1..M (src/calculate_observables.c:89-116,146-175). */
void or_energies(or_ctx *c, const double *psi, int M, double *E);

/* Observable codes (names of src/calculate_observables.c:48-56 plus
 * extensions). */
enum { OR_OBS_NONE = 0, OR_OBS_LZ = 1, OR_OBS_ENERGY = 2, OR_OBS_DIPOLE_R = 3,
       OR_OBS_DIPOLE_V = 4, OR_OBS_NORM = 5, OR_OBS_DIPOLE_A = 6 };
int or_obs_width(int code, int dim);
/* Values of the requested observables in the reference order (real parts,
 * src/calculate_observables.c:119-140).  alpha = vector potential (velocity
 * gauge) or zeros; Efield = E(t) used by the dipole_a extension. */
void or_observables(or_ctx *c, const double *psi, const int *codes, int ncodes,
                    const double *alpha, const double *Efield, double *out);

/* Mask extension: per-axis cos^(1/8) ramp on the outer `frac` of each
 * half-width.  frac <= 0 disables it. */
void or_set_mask(or_ctx *c, double frac);
void or_apply_mask(or_ctx *c, double *psi);

/* Imaginary-time driver (src/QuantumND_imag.c:169-240).
 * guess_mode 0: reference guess (radius+(i%3))*exp(-radius);
 * guess_mode 1: SplitMix64/Box-Muller complex Gaussian * exp(-|r|^2/20).
 * Outputs: psi_out [n_states][N] complex, E_out [n_states][2] (E, <H^2>),
 * Lz_out [n_states].  The global phase is removed as the reference does.
 * Returns 0 or 1 (Gram-Schmidt collapse). */
int or_imag_run(or_ctx *c, int n_states, double T, int guess_mode,
                uint64_t seed, double *psi_out, double *E_out, double *Lz_out);
/* The guesses alone (same generator), [n_states][N] complex. */
void or_initial_guess(or_ctx *c, int n_states, int guess_mode, uint64_t seed,
                      double *psi_out);
/* Steps the reference loop `for(t=0;t<T;t+=dt)` performs. */
int64_t or_imag_step_count(double T, double dt);

/* Real-time driver (src/QuantumND_real.c:166-228).  gauge 0 = length,
 * 1 = velocity.  psi: r-space in, r-space out.  Et, At: [ntime][s][dim]
 * stage tables; Eobs, Aobs: [ntime][dim] field at t_n for the observables.
 * Observables are recorded every step (the reference's dt_cum is never
 * reset), out = [ntime][width]. The mask (if set) is applied after every
 * step. */
int or_real_run(or_ctx *c, int gauge, double *psi, int64_t ntime,
                const double *Et, const double *At, const double *Eobs,
                const double *Aobs, const int *codes, int ncodes, double *out);

/* Field synthesis (src/parameters.c:899-949) for one time. env codes:
 * 0 constant, 1 trapz, 2 cos2, 3 cos4.  Returns 0, or 1 for a bad code. */
int or_field_vectors(int dim, int nmodes, const double *freq, const double *amp,
                     const double *ellip, const double *duration,
                     const double *cep, const double *delay, const int *env,
                     double t, double *A, double *E);
/* Extension: each mode's (x, y) field vectors rotated by angle[j] radians in
 * the xy plane (the reference fixes the major axis on x,
 * src/parameters.c:942-947).  angle == NULL is or_field_vectors. */
int or_field_vectors_angle(int dim, int nmodes, const double *freq, const double *amp,
                           const double *ellip, const double *duration,
                           const double *cep, const double *delay, const int *env,
                           const double *angle, double t, double *A, double *E);
/* Stage tables (src/integrators.c:74-129) [ntime][s][dim]. */
int or_field_tables(int dim, int nmodes, const double *freq, const double *amp,
                    const double *ellip, const double *duration,
                    const double *cep, const double *delay, const int *env,
                    const double *as_scaled, int s, const double *time,
                    int64_t ntime, double *A, double *E);
int or_field_tables_angle(int dim, int nmodes, const double *freq, const double *amp,
                          const double *ellip, const double *duration,
                          const double *cep, const double *delay, const int *env,
                          const double *angle, const double *as_scaled, int s,
                          const double *time, int64_t ntime, double *A, double *E);

/* Final observables (src/io_functions.c:136-287): mask, FFT to unitary
 * p-space, PED [Ne] and the PMD selection (density, phase) pairs in native
 * FFT order.  psi is overwritten by the masked p-space state.  Returns the
 * number of selected points; pmd may be NULL to count only. */
int64_t or_final_observables(or_ctx *c, double *psi, const double *R_mask,
                             const double *pr_max, int Ne, double de,
                             double sigma, double *ped, double *pmd,
                             int64_t *npr);

/* SplitMix64 stream -> standard normals (Box-Muller), n values. */
void or_gaussian(uint64_t seed, int64_t n, double *out);

#ifdef __cplusplus
}
#endif
#endif
