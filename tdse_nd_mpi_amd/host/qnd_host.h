/*
 * qnd_host.h -- host side of the QuantumND drop-in: parameter files, KEY VALUE
 * overrides, laser field synthesis and the output/input file formats.
 *
 * Behaviour (accepted keys, argument counts, defaults, messages, exit status
 * 1 on every error) follows the reference's src/parameters.c and
 * src/io_functions.c for a single process; the GPU work goes through libqnd
 * (include/qnd.h).
 */
#ifndef QND_HOST_H
#define QND_HOST_H

#include <stdint.h>
#include <stdio.h>

#define MAX_INTEGRATOR_CHAR 10   /* src/parameters.h:13 */
#define MAX_POTENTIAL_CHAR 20
#define MAX_OUTPUT_FILES_CHAR 30
#define MAX_OBSERVABLE_CHAR 32
#define MAX_GAUGE_CHAR 10
#define MAX_ENVELOPE_CHAR 10
#define MAX_PARAM_NAME_LEN 128

/* src/parameters.h:27-64 plus the extension keys */
typedef struct {
    int dim;
    int64_t *Nr_global;
    double *dr;
    int N_states;
    int *j_states;
    double *coeffs_states;        /* [N_states][2] (re, im) */
    double dt;
    double Tintegration;
    char integrator[MAX_INTEGRATOR_CHAR];
    char output_imag[MAX_OUTPUT_FILES_CHAR];
    char potential[MAX_POTENTIAL_CHAR];
    double cap[2];
    double *time;
    int64_t N_time;
    double dt_save;
    int N_observables;
    char **observables;
    double *pr_max;
    double R_mask[2];
    int Ne;
    double de;
    /* extensions (keys absent from the reference; defaults reproduce it) */
    int guess_mode;       /* GUESS reference|gaussian */
    uint64_t seed;        /* SEED */
    double softcore;      /* SOFTCORE a: V=-../sqrt(r^2+a^2); <0 = reference */
    int precision;        /* PRECISION c128|c64|c64f64 (QND_C128, QND_C64, QND_C64F64) */
    double mask_frac;     /* MASK f: cos^(1/8) ramp on the outer f of each half-width */
    int chunk;            /* STEPS_PER_CALL: real-time steps per library call */
} Parameters;

/* src/parameters.h:66-87 */
typedef struct {
    int N_modes;
    char gauge[MAX_GAUGE_CHAR];
    double *intensity, *lambda, *cep, *ellipticity, *duration, *delay;
    char **envelope;
    double *frequency, *amplitude;
    double *E, *A;        /* [N_time][s][dim] stage tables */
} Field;

/* Abort the program as the reference's MPI_Abort(comm,1) does. */
void qnd_abort(void);

/* parameters (src/parameters.c) */
void Initialize_parameters_imag(const char *filename, Parameters *P, int argc, char *argv[]);
void Initialize_parameters_real(const char *filename, Parameters *P, Parameters *P_imag, Field *F,
                                int argc, char *argv[]);
void Finalize_parameters(Parameters *P);
void Finalize_field_vectors(Field *F);
int Expected_arguments(FILE *file_ptr, const char *parameter_name, int expected_number,
                       const char *KnownKeys[]);
void parse_complex(const char *str, double *re, double *im);
int Find_dimension_in_file(const char *filename);

/* Field_vectors (src/parameters.c:899-949): A and E at time t */
void Field_vectors(const Parameters *P, const Field *F, double t, double *A, double *E);
/* Initialize_field_vectors (src/integrators.c:74-129), writes the two field
 * files; as_scaled from the integrator of s stages. */
void Initialize_field_vectors(const Parameters *P, Field *F, const double *as_scaled, int s);

/* integrator coefficients (src/integrators.c:13-71); -1 for an unknown name */
int integrator_coefficients(const char *name, double dt, double *as, double *as_scaled);

/* SplitMix64/Box-Muller complex Gaussian guesses (extension) */
void gaussian_stream(uint64_t seed, int64_t n, double *out);

/* I/O (src/io_functions.c) */
void softcore_side_file(const char *output_imag, char *buf, size_t len);
void Write_output_imag(const char *filename, double **Psi, double **Ej, const Parameters *P, int64_t N);
void Write_observables(FILE *f, double time, const double *values, int nvalues);
void Write_wavefunction_to_file(const char *filename, const double *Psi, const Parameters *P, int64_t N);
/* Read_wavefunction (src/io_functions.c:292-459): centred embedding and
 * superposition into Psi (zeroed by the caller), single process. */
void Read_wavefunction(const Parameters *P, const Parameters *P_imag, double *Psi);
void Write_PED(const double *S, int Ne, double de);
void Write_PMD(const double *psi_p, const Parameters *P);

/* observable name -> code and width (src/calculate_observables.c:48-56) */
int observable_code(const char *name);
int observable_width(const char *name, int dim);

double wall_seconds(void);

#endif
