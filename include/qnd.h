/*
 * qnd.h -- C ABI of libqnd, the MI355X (gfx950) split-operator TDSE hot path
 * of QuantumND.
 *
 * The reference has no FFI: its hot path is a set of C functions called from
 * the two drivers (src/QuantumND_imag.c, src/QuantumND_real.c) on
 * MPI-distributed host arrays.  libqnd replaces that in-process seam with
 * plain-pointer entry points over one GPU: the whole N-D grid lives in HBM,
 * tables are analytic on the device, and many steps run per call.  Each
 * declaration names the reference interface it replaces.
 *
 * Conventions
 *  - Every entry returns 0 on success and a non-zero QND_E* code on error;
 *    qnd_last_error() gives the message.  The library never exits.
 *  - One host thread per context; each context owns two HIP streams (the
 *    second overlaps the adjoint's recompute with its costate sweep).  Calls
 *    are stream-ordered; host-visible outputs are complete on return.
 *  - Complex host arrays are interleaved FP64 (re, im), row-major, axis 0
 *    slowest: the reference layout (src/meshgrid.c:205-211).
 *  - Field tables use the reference layout [step][stage][dim] in FP64
 *    (src/integrators.c:74-129).
 *  - A context holds n_slots state vectors ("slots").  States are in
 *    position representation at the interface; the internal representation
 *    between calls is the library's business.
 */
#ifndef QND_H
#define QND_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qnd_ctx qnd_ctx;

/* precision of the propagated state */
enum { QND_C128 = 0,     /* complex128 state, FP64 everywhere */
       QND_C64 = 1,      /* complex64 state, FP32 phases from FP64-reduced
                            angles, FP64 reductions */
       QND_C64F64 = 2 }; /* complex64 state in memory, FP64 arithmetic in
                            the passes (each value rounded once per pass,
                            at its store); phases as for QND_C64 */

enum { QND_GAUGE_LENGTH = 0, QND_GAUGE_VELOCITY = 1 };

/* Observable codes.  1-4 are the reference's names
 * (src/calculate_observables.c:48-56); 5-6 are extensions. */
enum { QND_OBS_NONE = 0, QND_OBS_LZ = 1, QND_OBS_ENERGY = 2,
       QND_OBS_DIPOLE_R = 3, QND_OBS_DIPOLE_V = 4,
       QND_OBS_NORM = 5,      /* <psi|psi> */
       QND_OBS_DIPOLE_A = 6   /* -<grad V> - E(t)<psi|psi> (Ehrenfest) */ };

/* error codes */
enum { QND_OK = 0, QND_EINVAL = 1, QND_EHIP = 2, QND_EUNKNOWN_NAME = 3,
       QND_EGS_COLLAPSE = 4, QND_EUNSUPPORTED = 5, QND_ESTATE = 6 };

/* Create a context for a dim-D grid of nr[j] points spaced dr[j] with
 * n_slots state vectors.  Replaces the grid part of Initialize_meshgrid
 * (src/meshgrid.c:17-179) and the fftw_mpi plan creation
 * (src/meshgrid.c:159-167): axis grids r, p (FFT order), dNr, dNp and
 * fft_scaling are defined exactly as there; FFT plans are this library's
 * own Stockham passes.  No device is touched before this call. */
int qnd_create(qnd_ctx **ctx, int dim, const int64_t *nr, const double *dr,
               int precision, int n_slots);

/* Finalize_meshgrid (src/meshgrid.c:184-199) + Finalize_integrator. */
void qnd_destroy(qnd_ctx *ctx);

/* Message of the last error of this context (or of a failed qnd_create when
 * ctx is NULL). */
const char *qnd_last_error(const qnd_ctx *ctx);

/* Potential and complex absorbing potential: the Vpot/Vcap tables of
 * src/meshgrid.c:123-154, evaluated analytically on the device.  softcore < 0
 * keeps the reference constant; softcore >= 0 uses V = -(..)/sqrt(r^2 +
 * softcore^2) (extension).  cap = {strength, width} as CAP in params.real
 * (already |.|-ed, src/parameters.c:526-534); NULL means no CAP.
 * Unknown name -> QND_EUNKNOWN_NAME ("Unknown potential: <name>"). */
int qnd_set_potential(qnd_ctx *ctx, const char *name, double softcore,
                      const double *cap);

/* Initialize_integrator (src/integrators.c:13-71): Verlet, BM4, BM6, RKN4. */
int qnd_set_integrator(qnd_ctx *ctx, const char *name, double dt);

/* Extension: cos^(1/8) absorbing mask on the outer `frac` of each half-width,
 * applied after every real-time step.  frac <= 0 disables it. */
int qnd_set_mask(qnd_ctx *ctx, double frac);

/* State transfer.  psi: host interleaved FP64 complex, position space. */
int qnd_upload_state(qnd_ctx *ctx, int slot, const double *psi);
int qnd_download_state(qnd_ctx *ctx, int slot, double *psi);

/* Unnormalised N-D DFT of slot in place, sign -1 (forward) or +1
 * (backward): the fftw_execute(plan_forward / plan_backward) of
 * src/meshgrid.c:166-167 on this library's own Stockham passes.  The slot is
 * taken in position space and left holding the transform. */
int qnd_transform(qnd_ctx *ctx, int slot, int sign);

/* nsteps imaginary-time steps of slot, each followed (gram_schmidt != 0) by
 * modified Gram-Schmidt against slots 0..slot-1.  Replaces the loop body
 * src/QuantumND_imag.c:182-185: Integration_step_imag
 * (src/integrators.c:135-163) + Gram_Schmidt_orthonormalization
 * (src/QuantumND_imag.c:265-311).  QND_EGS_COLLAPSE when the norm falls
 * below 1e-12. */
int qnd_imag_steps(qnd_ctx *ctx, int slot, int64_t nsteps, int gram_schmidt);

/* Standalone Gram_Schmidt_orthonormalization of slot against 0..slot-1. */
int qnd_gram_schmidt(qnd_ctx *ctx, int slot);

/* Calculation_global_energies (src/calculate_observables.c:89-116):
 * E[m-1] = Re <psi|H^m|psi>, m = 1..M. */
int qnd_energy_moments(qnd_ctx *ctx, int slot, int M, double *E);

/* Calculation_global_observables (src/calculate_observables.c:119-140) on
 * slot: values of the requested codes, in order, real parts.  alpha = A(t)
 * for the velocity gauge (NULL = 0); Efield = E(t) for dipole_a (NULL = 0). */
int qnd_observables(qnd_ctx *ctx, int slot, const int *codes, int ncodes,
                    const double *alpha, const double *Efield, double *out);

/* Allocation-free variant of qnd_observables, with no counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Computes the fused position-space codes QND_OBS_NORM, QND_OBS_DIPOLE_R and
 * QND_OBS_DIPOLE_A (others: QND_EUNSUPPORTED) of the state psi_dev, a device
 * array of N elements of the context's precision (float2 for QND_C64 and
 * QND_C64F64, double2 for QND_C128,
 * row-major, position space), into out_dev, a device array of the codes'
 * widths (as qnd_observables).  Efield (host, dim values or NULL) is E(t) of
 * dipole_a.  It calls no allocating routine: its reduction partials and sums
 * live in a fixed device array of QND_NOALLOC_ROWS (8192) partial rows per
 * GPU, held until the process ends; a grid needing more rows is refused
 * (QND_EINVAL).  The call is ordered on the context's stream and returns
 * without waiting: out_dev is complete once the stream reaches it (e.g.
 * after qnd_synchronize).  Calls that overlap on one GPU share that array
 * and race; the caller must order them. */
int qnd_observables_noalloc(qnd_ctx *ctx, const void *psi_dev, const int *codes,
                            int ncodes, const double *Efield, double *out_dev);

/* Field tables of Initialize_field_vectors (src/integrators.c:74-129):
 * Et, At [ntime][s][dim] at the stage times; Eobs, Aobs [ntime][dim] at the
 * step times t_n (Field_vectors of src/QuantumND_real.c:173).  Copied. */
int qnd_set_field(qnd_ctx *ctx, int64_t ntime, const double *Et,
                  const double *At, const double *Eobs, const double *Aobs);

/* Real-time steps step0 .. step0+nsteps-1 of slot in the given gauge.
 * Replaces the loop body of src/QuantumND_real.c:169-186 (velocity) and
 * :195-223 (length): Calculation_global_observables every step (the
 * reference's dt_cum is never reset) followed by Integration_step_velocity
 * (src/integrators.c:201-232) or Integration_step_length
 * (src/integrators.c:170-194), then the mask extension.
 * obs_out: [nsteps][width] host array (NULL when ncodes == 0), width = sum of
 * widths of the codes (1 for Lz/energy/norm, dim for dipoles). */
int qnd_real_steps(qnd_ctx *ctx, int slot, int gauge, int64_t step0,
                   int64_t nsteps, const int *codes, int ncodes,
                   double *obs_out);

/* Adjoint gradients (no reference counterpart: the reference has no
 * differentiation).  qnd_forward_checkpointed runs nsteps length-gauge steps
 * of slot from step 0 (dim >= 2), recording per step the dipole
 * acceleration and the norm (obs_out [nsteps][dim+1], may be NULL) and the
 * norm after the final owed mask; the state is checkpointed on the device
 * every `every` steps.  qnd_adjoint then back-propagates
 *     L = f(a_0 .. a_{nsteps-1}) + g(norm_final)
 * given dL_da [nsteps][dim] and dL_dnorm: the costate runs backwards through
 * the conjugated passes, each checkpoint segment is recomputed once, and it
 * returns dL/dEt [nsteps][s][dim] (the stage fields of qnd_set_field),
 * dL/dEobs [nsteps][dim] (the E(t_n) inside a_n) and dL/d(soft-core length)
 * (the initial state held fixed).  nsteps is the row count of the caller's
 * arrays and must equal the forward run's (else QND_EINVAL).  qnd_adjoint
 * always differentiates the context's last checkpointed forward run; changing
 * the potential, integrator, mask or field between the two calls makes it
 * fail with QND_ESTATE. */
int qnd_forward_checkpointed(qnd_ctx *ctx, int slot, int64_t nsteps,
                             int64_t every, double *obs_out,
                             double *norm_final);
int qnd_adjoint(qnd_ctx *ctx, int64_t nsteps, const double *dL_da,
                double dL_dnorm, double *dL_dEt, double *dL_dEobs,
                double *dL_dsoft);

/* Launch tuning (no reference counterpart: FFTW's planner plays this role,
 * src/meshgrid.c:166-167 with FFTW_MEASURE).  vpt = values per thread of a
 * power-of-two line (4 or 8), tw_* = lines per workgroup for contiguous and
 * strided passes; 0 = automatic.  Any value gives the same results up to
 * rounding; the library clamps them to what a line and the LDS allow.
 * tdse_nd_mpi_amd.tuning times candidates on the live state. */
int qnd_set_tuning(qnd_ctx *ctx, int vpt, int tw_contig, int tw_strided);

/* Batched real time (no reference counterpart: the reference runs one
 * field per process; this serves CEP / intensity sweeps).  Instance b of a
 * batch is slot slot0+b, propagated under its own field:
 * Et, At [nb][ntime][s][dim] at the stage times and Eobs [nb][ntime][dim] at
 * t_n, each instance's rows in the layout of qnd_set_field.  Copied to the
 * device.  nb must not exceed n_slots. */
int qnd_set_field_batch(qnd_ctx *ctx, int nb, int64_t ntime, const double *Et,
                        const double *At, const double *Eobs);

/* Field synthesis on the device.  Field_vectors (src/parameters.c:899-949)
 * at the stage times of Initialize_field_vectors (src/integrators.c:74-129)
 * and at t_n = n DT (Q2), for nb instances of nmodes modes each: modes is a
 * host array [nb][nmodes]; instance b's tables are those qnd_set_field_batch
 * would receive from Field_vectors with its modes, formed by a kernel
 * directly in the batch tables (no host tables, no copy).  A mode is the
 * reference's laser mode (parameters.c:737-738: F0 = 5.33803e-9 sqrt(I),
 * omega = 45.5637639947958 / lambda; A = F0 f (cos, xi sin)(omega t + phi) /
 * omega on (x, y), E = -dA/dt with the envelope's analytic derivative), plus
 * `angle`, an extension that rotates the mode's (x, y) vector in the xy plane
 * (0 reproduces the reference).  Uses the context's integrator and DT.  With
 * nb == 1 the single-instance tables of qnd_set_field (including A(t_n)) are
 * set too.  Errors: QND_ESTATE without an integrator; QND_EINVAL for nb
 * outside [1, n_slots], nmodes < 1, an unknown envelope, a wavelength <= 0,
 * an intensity < 0 or a duration <= 0. */
#define QND_ENV_CONSTANT 0
#define QND_ENV_TRAPZ 1
#define QND_ENV_COS2 2
#define QND_ENV_COS4 3
typedef struct qnd_mode {
    double intensity;   /* W/cm^2 */
    double wavelength;  /* nm */
    double ellipticity; /* amplitude ratio of the y component */
    double cep;         /* rad */
    double delay;       /* a.u. of time */
    double duration;    /* envelope length tau, a.u. of time */
    double angle;       /* rad, rotation of the (x, y) vector (extension) */
    int envelope;       /* QND_ENV_* */
    int pad_;
} qnd_mode;
int qnd_set_field_modes(qnd_ctx *ctx, int nb, int64_t ntime, int nmodes,
                        const qnd_mode *modes);

/* Copy the batch field tables to the host (any pointer may be NULL):
 * Et, At [nb][ntime][s][dim], Eobs [nb][ntime][dim] with nb and ntime those
 * of the last qnd_set_field_batch / qnd_set_field_modes call (no reference
 * counterpart: the executables write the single-instance tables to
 * electric_field.real / vector_potential.real, src/integrators.c:103-127). */
int qnd_get_field_batch(qnd_ctx *ctx, double *Et, double *At, double *Eobs);

/* qnd_real_steps for the nb slots slot0 .. slot0+nb-1 at once: every pass of
 * a step is one launch over all instances (the batch is an extra,
 * untransformed leading axis), each instance reading its own field rows.
 * Recordable codes: dipole_r, dipole_a, norm.  obs_out: [nb][nsteps][width].
 * With nb == 1 this is qnd_real_steps on the qnd_set_field tables. */
int qnd_real_steps_batch(qnd_ctx *ctx, int slot0, int nb, int gauge,
                         int64_t step0, int64_t nsteps, const int *codes,
                         int ncodes, double *obs_out);

/* Several grids stepped together (no reference counterpart: the reference
 * runs one grid per process; this serves sweeps over uneven grid sizes).
 * Each member context ctxs[k] (2-D, one precision, one integrator and DT, at
 * most 4 members, each with its own stream) propagates its batch of slots
 * 0 .. nb_k - 1 under its qnd_set_field_batch tables, as
 * qnd_real_steps_batch would.  Length gauge; recordable codes dipole_r,
 * dipole_a, norm; obs_out[k]: [nb_k][nsteps][width] host array of member k.
 * The steps after the first advance a device step counter, so one step's
 * launch sequence serves them all.  use_graph bit 0: that sequence is
 * captured as a HIP graph and replayed; bit 1 (the branch form): each pass
 * is each member's own kernel on its own stream, forked from and joined to
 * the first member's stream every step (one graph launch then runs up to 8
 * steps of all members); without bit 1 every pass of a step is one kernel
 * launch over all members' instances. */
int qnd_group_real_steps(qnd_ctx *const *ctxs, int nctx, int gauge,
                         int64_t step0, int64_t nsteps, const int *codes,
                         int ncodes, double *const *obs_out, int use_graph);

/* Write_final_observables (src/io_functions.c:136-287) numerics: radial
 * sin^2 mask (R_mask), FFT to unitary p-space; PED[Ne] (sigma = de) and the
 * masked p-space state psi_p (host, interleaved, native FFT order) from
 * which the PMD records are selected.  Either output may be NULL. */
int qnd_final_observables(qnd_ctx *ctx, int slot, const double *R_mask,
                          const double *pr_max, int Ne, double de,
                          double *ped, double *psi_p);

/* Host-side timing on the context stream: record event idx (0..15), and the
 * elapsed milliseconds between two recorded events (synchronizes). */
int qnd_event_record(qnd_ctx *ctx, int idx);
int qnd_event_elapsed(qnd_ctx *ctx, int i0, int i1, float *ms);
int qnd_synchronize(qnd_ctx *ctx);

/* Per-kernel-class profiling: when enabled, every launch of a pass class is
 * bracketed by events; qnd_profile_read returns the count and total ms per
 * class (classes: 0 potential/axis-0 pass, 1 kinetic/last-axis pass,
 * 2 plain FFT pass, 3 small reduction kernels and record passes, 4 the two
 * passes of a merged step boundary). */
int qnd_profile_enable(qnd_ctx *ctx, int enable);
int qnd_profile_read(qnd_ctx *ctx, int64_t *count, double *ms, int nclass);

/* Library version string. */
const char *qnd_version(void);

#ifdef __cplusplus
}
#endif
#endif /* QND_H */
