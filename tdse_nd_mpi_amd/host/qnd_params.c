/*
 * qnd_params.c -- params.imag / params.real parsing, KEY VALUE overrides and
 * the laser field, with the semantics of src/parameters.c (single process:
 * the reference's rank-0 branch, without the MPI broadcasts) and
 * src/integrators.c:13-129.
 *
 * Extension keys (not in the reference; files that do not use them parse
 * exactly as in the reference):
 *   params.imag: GUESS reference|gaussian, SEED n, SOFTCORE a, PRECISION c128|c64|c64f64
 *   params.real: MASK f, PRECISION c128|c64|c64f64, STEPS_PER_CALL n
 */
#include "qnd_host.h"

#include <math.h>
#include <stdbool.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

void qnd_abort(void)
{
    fflush(stdout);
    fflush(stderr);
    exit(1);
}

double wall_seconds(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + ts.tv_nsec * 1e-9;
}

static bool is_known_key(const char *tok, const char **keys)
{
    for (int k = 0; keys[k] != NULL; k++)
        if (strcmp(tok, keys[k]) == 0) return true;
    return false;
}

/* src/parameters.c:1080-1130: exactly expected_number tokens must follow the
 * key before the next known key or EOF; the file position is restored. */
int Expected_arguments(FILE *fp, const char *name, int expected, const char *KnownKeys[])
{
    char buf[MAX_PARAM_NAME_LEN];
    long start = ftell(fp);
    int count = 0;
    while (count < expected) {
        if (fscanf(fp, "%127s", buf) != 1) {
            fprintf(stderr,
                    "Rank 0 ERROR reading file: Parameter '%s' expected %d argument(s), but found only %d "
                    "(reached EOF).\n",
                    name, expected, count);
            qnd_abort();
        }
        if (is_known_key(buf, KnownKeys)) {
            fprintf(stderr,
                    "Rank 0 ERROR reading file: Parameter '%s' expected %d argument(s), but found only %d "
                    "(hit new parameter '%s').\n",
                    name, expected, count, buf);
            qnd_abort();
        }
        count++;
    }
    if (fscanf(fp, "%127s", buf) == 1 && !is_known_key(buf, KnownKeys)) {
        fprintf(stderr,
                "Rank 0 ERROR reading file: Parameter '%s' expected exactly %d argument(s), but found extra, "
                "unrecognized token '%s'.\n",
                name, expected, buf);
        qnd_abort();
    }
    if (fseek(fp, start, SEEK_SET) != 0) {
        perror("Error rewinding file pointer in Expected_arguments");
        qnd_abort();
    }
    return 0;
}

/* src/parameters.c:1058-1077: "a+bj", "a-bj", "bj", "a" */
void parse_complex(const char *str, double *re, double *im)
{
    double a = 0., b = 0.;
    const char *plus = strchr(str, '+');
    const char *minus = strrchr(str, '-');
    if (plus && strchr(str, 'j')) {
        sscanf(str, "%lf+%lfj", &a, &b);
    } else if (minus && minus != str && strchr(str, 'j')) {
        sscanf(str, "%lf-%lfj", &a, &b);
        b = -b;
    } else if (strchr(str, 'j')) {
        sscanf(str, "%lfj", &b);
        a = 0.;
    } else {
        sscanf(str, "%lf", &a);
        b = 0.;
    }
    *re = a;
    *im = b;
}

/* src/parameters.c:998-1029 */
int Find_dimension_in_file(const char *filename)
{
    int dim = 0;
    FILE *f = fopen(filename, "r");
    if (!f) {
        perror("Error opening parameter file");
        qnd_abort();
    }
    char key[32];
    while (fscanf(f, "%31s", key) != EOF) {
        if (strcmp(key, "DIMENSION") == 0) {
            if (fscanf(f, "%d", &dim) != 1) {
                perror("The dimension is undetermined");
                qnd_abort();
            }
            break;
        }
    }
    fclose(f);
    if (dim == 0) {
        printf("\nThe dimension is not specified\n");
        qnd_abort();
    }
    return dim;
}

/* src/parameters.c:1032-1055 */
static void allocate_dim_arrays(Parameters *P, int dim)
{
    P->dim = dim;
    P->Nr_global = (int64_t *)malloc(dim * sizeof(int64_t));
    P->dr = (double *)malloc(dim * sizeof(double));
    P->pr_max = (double *)malloc(dim * sizeof(double));
    if (!P->Nr_global || !P->dr || !P->pr_max) {
        fprintf(stderr, "ERROR: memory allocation failed for Nr_global or dr.\n");
        qnd_abort();
    }
    for (int j = 0; j < dim; j++) {
        P->Nr_global[j] = 256;
        P->dr[j] = .1;
        P->pr_max[j] = 1.;
    }
}

static void extension_defaults(Parameters *P)
{
    P->guess_mode = 0;
    P->seed = 0;
    P->softcore = -1.0;
    P->precision = 0;
    P->mask_frac = 0.0;
    P->chunk = 256;
}

static int parse_precision(const char *s)
{
    if (strcmp(s, "c128") == 0) return 0;
    if (strcmp(s, "c64") == 0) return 1;
    if (strcmp(s, "c64f64") == 0) return 2;
    printf("\nUnknown precision: %s (c128, c64 or c64f64)\n", s);
    qnd_abort();
    return 0;
}

static bool is_pow2(int64_t n) { return n == 0 || (n & (n - 1)) == 0; }

/* ------------------------------------------------------------------ */
/* params.imag (src/parameters.c:11-212)                                */
/* ------------------------------------------------------------------ */
void Initialize_parameters_imag(const char *filename, Parameters *P, int argc, char *argv[])
{
    static const char *Known[] = {"OUTPUT_IMAG", "DIMENSION", "DR", "NR", "N_STATES", "POTENTIAL",
                                  "INTEGRATOR", "DT", "TINTEGRATION",
                                  "GUESS", "SEED", "SOFTCORE", "PRECISION", NULL};
    char key[32], val[64];
    int j;
    snprintf(P->output_imag, MAX_OUTPUT_FILES_CHAR, "output.imag");
    P->N_states = 1;
    snprintf(P->potential, MAX_POTENTIAL_CHAR, "H_3D");
    snprintf(P->integrator, MAX_INTEGRATOR_CHAR, "BM4");
    P->dt = .1;
    P->Tintegration = 100.;
    extension_defaults(P);

    int dim = Find_dimension_in_file(filename);
    allocate_dim_arrays(P, dim);

    FILE *f = fopen(filename, "r");
    if (!f) {
        perror("Error opening parameter file");
        qnd_abort();
    }
    while (fscanf(f, "%31s", key) != EOF) {
        if (strcmp(key, "DIMENSION") == 0) {
            Expected_arguments(f, "DIMENSION", 1, Known);
            if (fscanf(f, "%31s", key) != 1) qnd_abort();
        } else if (strcmp(key, "NR") == 0) {
            Expected_arguments(f, "NR", P->dim, Known);
            for (j = 0; j < P->dim; j++) {
                long long v;
                if (fscanf(f, "%lld", &v) != 1) {
                    perror("The number of grid points is not completely determined");
                    qnd_abort();
                }
                P->Nr_global[j] = v;
                if (!is_pow2(P->Nr_global[j])) {
                    printf("\nThe number of points on the grid must be a power of 2\n");
                    perror("Error with grid points number");
                    qnd_abort();
                }
            }
        } else if (strcmp(key, "DR") == 0) {
            Expected_arguments(f, "DR", P->dim, Known);
            for (j = 0; j < P->dim; j++)
                if (fscanf(f, "%lf", &P->dr[j]) != 1) {
                    perror("The grid spacing is not completely determined");
                    qnd_abort();
                }
        } else if (strcmp(key, "N_STATES") == 0) {
            Expected_arguments(f, "N_STATES", 1, Known);
            if (fscanf(f, "%d", &P->N_states) != 1) {
                perror("The number of states is undetermined");
                qnd_abort();
            }
        } else if (strcmp(key, "DT") == 0) {
            Expected_arguments(f, "DT", 1, Known);
            if (fscanf(f, "%lf", &P->dt) != 1) {
                perror("The time step is undetermined");
                qnd_abort();
            }
        } else if (strcmp(key, "TINTEGRATION") == 0) {
            Expected_arguments(f, "TINTEGRATION", 1, Known);
            if (fscanf(f, "%lf", &P->Tintegration) != 1) {
                perror("The time integration is undetermined");
                qnd_abort();
            }
        } else if (strcmp(key, "POTENTIAL") == 0) {
            Expected_arguments(f, "POTENTIAL", 1, Known);
            if (fscanf(f, "%19s", P->potential) != 1) {
                perror("The potential name is undetermined");
                qnd_abort();
            }
        } else if (strcmp(key, "INTEGRATOR") == 0) {
            Expected_arguments(f, "INTEGRATOR", 1, Known);
            if (fscanf(f, "%9s", P->integrator) != 1) {
                perror("The integrator name is undetermined");
                qnd_abort();
            }
        } else if (strcmp(key, "OUTPUT_IMAG") == 0) {
            Expected_arguments(f, "OUTPUT_IMAG", 1, Known);
            if (fscanf(f, "%29s", P->output_imag) != 1) {
                perror("The name of OUTPUT_IMAG is undetermined");
                qnd_abort();
            }
        } else if (strcmp(key, "GUESS") == 0) {
            Expected_arguments(f, "GUESS", 1, Known);
            if (fscanf(f, "%63s", val) != 1) qnd_abort();
            if (strcmp(val, "reference") == 0) P->guess_mode = 0;
            else if (strcmp(val, "gaussian") == 0) P->guess_mode = 1;
            else {
                printf("\nUnknown guess: %s\n", val);
                qnd_abort();
            }
        } else if (strcmp(key, "SEED") == 0) {
            Expected_arguments(f, "SEED", 1, Known);
            unsigned long long s;
            if (fscanf(f, "%llu", &s) != 1) qnd_abort();
            P->seed = s;
        } else if (strcmp(key, "SOFTCORE") == 0) {
            Expected_arguments(f, "SOFTCORE", 1, Known);
            if (fscanf(f, "%lf", &P->softcore) != 1) qnd_abort();
        } else if (strcmp(key, "PRECISION") == 0) {
            Expected_arguments(f, "PRECISION", 1, Known);
            if (fscanf(f, "%63s", val) != 1) qnd_abort();
            P->precision = parse_precision(val);
        } else {
            printf("\nUnknown parameter: %s\n", key);
            perror("Error with params.txt");
            qnd_abort();
        }
    }
    fclose(f);
    if (P->dim == 0 || P->Nr_global[0] <= 0 || P->N_states <= 0) {
        fprintf(stderr, "ERROR: Could not read or interpret the file %s properly.\n", filename);
        fprintf(stderr, "\n--- FATAL ERROR: Parameter file processing failed. Aborting job. ---\n");
        qnd_abort();
    }
    /* overrides (src/parameters.c:157-184) */
    for (j = 2; j < argc; j += 2) {
        if (j + 1 < argc) {
            const char *k = argv[j], *v = argv[j + 1];
            if (strcmp(k, "DT") == 0) P->dt = atof(v);
            else if (strcmp(k, "TINTEGRATION") == 0) P->Tintegration = atof(v);
            else if (strcmp(k, "OUTPUT_IMAG") == 0) {
                strncpy(P->output_imag, v, MAX_OUTPUT_FILES_CHAR);
                P->output_imag[MAX_OUTPUT_FILES_CHAR - 1] = '\0';
            } else {
                fprintf(stderr, "Rank 0 ERROR: Unrecognized parameter key '%s'. Check spelling.\n", k);
                qnd_abort();
            }
        } else {
            fprintf(stderr, "Rank 0 ERROR: Parameter key '%s' found without a corresponding value.\n", argv[j]);
            qnd_abort();
        }
    }
    printf("  DIMENSION\t\t%d\n", P->dim);
    printf("  NR\t\t");
    for (j = 0; j < P->dim; j++) printf("\t%lld", (long long)P->Nr_global[j]);
    printf("\n  DR\t\t");
    for (j = 0; j < P->dim; j++) printf("\t%16.15f", P->dr[j]);
    printf("\n");
    printf("  N_STATES\t\t%d\n", P->N_states);
    printf("  POTENTIAL\t\t%s\n", P->potential);
    printf("  INTEGRATOR\t\t%s\n", P->integrator);
    printf("  DT\t\t\t%16.15f\n", P->dt);
    printf("  TINTEGRATION\t\t%16.15f\n", P->Tintegration);
    printf("  OUTPUT_IMAG\t\t%s\n", P->output_imag);
    fflush(stdout);
}

/* ------------------------------------------------------------------ */
/* params.real (src/parameters.c:215-896)                               */
/* ------------------------------------------------------------------ */
static void read_doubles(FILE *f, double *dst, int n, const char *msg, int *err)
{
    for (int j = 0; j < n; j++)
        if (fscanf(f, "%lf", &dst[j]) != 1) {
            perror(msg);
            *err = 1;
        }
}

void Initialize_parameters_real(const char *filename, Parameters *P, Parameters *P_imag, Field *F,
                                int argc, char *argv[])
{
    static const char *Known[] = {"OUTPUT_IMAG", "NR", "N_STATES", "J_STATES", "COEFFS_STATES", "CAP",
                                  "GAUGE", "N_MODES", "INTENSITY", "LAMBDA", "ELLIPTICITY", "CEP",
                                  "ENVELOPE", "DURATION", "DELAY", "INTEGRATOR", "DT", "TINTEGRATION",
                                  "DT_SAVE", "N_OBSERVABLES", "OBSERVABLES", "P_MAX", "R_MASK", "DE", "NE",
                                  "MASK", "PRECISION", "STEPS_PER_CALL", NULL};
    char key[32], val[64];
    int j, local_error = 0;

    P->N_states = 1;
    snprintf(P->integrator, MAX_INTEGRATOR_CHAR, "BM4");
    P->dt = .1;
    P->Tintegration = 220.6380;
    P->cap[0] = P->cap[1] = 0.;
    F->N_modes = 1;
    snprintf(F->gauge, MAX_GAUGE_CHAR, "velocity");
    P->dt_save = 1.;
    P->N_observables = 0;
    P->R_mask[0] = 20.;
    P->R_mask[1] = 30.;
    P->Ne = 100;
    P->de = .01;
    extension_defaults(P);
    extension_defaults(P_imag);

    /* pass 1: sizes and the imaginary-time output name (:261-330) */
    FILE *f = fopen(filename, "r");
    if (!f) {
        perror("Error opening parameter file");
        fprintf(stderr, "\n FATAL ERROR: Critical parameters missing. Aborting.\n");
        qnd_abort();
    }
    int output_found = 0, nstates_found = 0, nobs_found = 0, nmodes_found = 0;
    while (fscanf(f, "%31s", key) != EOF) {
        if (strcmp(key, "OUTPUT_IMAG") == 0) {
            Expected_arguments(f, "OUTPUT_IMAG", 1, Known);
            if (fscanf(f, "%29s", P->output_imag) != 1) {
                perror("The name of output file is undetermined");
                qnd_abort();
            }
            output_found = 1;
        } else if (strcmp(key, "N_STATES") == 0) {
            Expected_arguments(f, "N_STATES", 1, Known);
            if (fscanf(f, "%d", &P->N_states) != 1) {
                perror("The number of states is undetermined");
                qnd_abort();
            }
            nstates_found = 1;
        } else if (strcmp(key, "N_OBSERVABLES") == 0) {
            Expected_arguments(f, "N_OBSERVABLES", 1, Known);
            if (fscanf(f, "%d", &P->N_observables) != 1) {
                perror("The number of observables is undetermined");
                qnd_abort();
            }
            nobs_found = 1;
        } else if (strcmp(key, "N_MODES") == 0) {
            Expected_arguments(f, "N_MODES", 1, Known);
            if (fscanf(f, "%d", &F->N_modes) != 1) {
                perror("The number of field modes is undetermined");
                qnd_abort();
            }
            nmodes_found = 1;
        }
        if (output_found && nstates_found && nobs_found && nmodes_found) break;
    }
    fclose(f);
    if (!output_found) {
        perror("The output file is not specified");
        local_error = 1;
    }
    if (!nstates_found) printf("WARNING: The number of states is not specified\nDefault value: N_STATES=1\n");
    if (!nobs_found)
        printf("WARNING: The number of observables is not specified\nDefault value: N_OBSERVABLES=0\n");
    if (!nmodes_found) printf("WARNING: The number of field modes is not specified\nDefault value: N_MODES=1\n");
    if (local_error) {
        fprintf(stderr, "\n FATAL ERROR: Critical parameters missing. Aborting.\n");
        qnd_abort();
    }

    /* allocation and defaults (:346-408) */
    P->j_states = (int *)malloc((P->N_states > 0 ? P->N_states : 1) * sizeof(int));
    P->coeffs_states = (double *)malloc(2 * (P->N_states > 0 ? P->N_states : 1) * sizeof(double));
    P->observables = (char **)malloc((P->N_observables > 0 ? P->N_observables : 1) * sizeof(char *));
    int nm = F->N_modes > 0 ? F->N_modes : 1;
    F->envelope = (char **)malloc(nm * sizeof(char *));
    F->intensity = (double *)malloc(nm * sizeof(double));
    F->lambda = (double *)malloc(nm * sizeof(double));
    F->ellipticity = (double *)malloc(nm * sizeof(double));
    F->cep = (double *)malloc(nm * sizeof(double));
    F->delay = (double *)malloc(nm * sizeof(double));
    F->duration = (double *)malloc(nm * sizeof(double));
    F->frequency = (double *)malloc(nm * sizeof(double));
    F->amplitude = (double *)malloc(nm * sizeof(double));
    for (j = 0; j < P->N_observables; j++) {
        P->observables[j] = (char *)malloc(MAX_OBSERVABLE_CHAR);
        snprintf(P->observables[j], MAX_OBSERVABLE_CHAR, "none");
    }
    for (j = 0; j < P->N_states; j++) {
        P->j_states[j] = 1;
        P->coeffs_states[2 * j] = 1.;
        P->coeffs_states[2 * j + 1] = 0.;
    }
    for (j = 0; j < F->N_modes; j++) {
        F->envelope[j] = (char *)malloc(MAX_ENVELOPE_CHAR);
        snprintf(F->envelope[j], MAX_ENVELOPE_CHAR, "cos4");
        F->intensity[j] = 1.e14;
        F->lambda[j] = 800.;
        F->ellipticity[j] = 1.;
        F->cep[j] = 0.;
        F->delay[j] = 0.;
        F->duration[j] = 220.6380;
    }

    int dim = Find_dimension_in_file(P->output_imag);
    allocate_dim_arrays(P, dim);
    allocate_dim_arrays(P_imag, dim);
    P_imag->N_states = 1;

    /* header of OUTPUT_IMAG (:416-466) */
    local_error = 0;
    f = fopen(P->output_imag, "r");
    if (f) {
        while (fscanf(f, "%31s", key) != EOF) {
            if (strcmp(key, "PSI") == 0) break;
            else if (strcmp(key, "DIMENSION") == 0) {
                if (fscanf(f, "%31s", key) != 1) local_error = 1;
            } else if (strcmp(key, "N_STATES") == 0) {
                if (fscanf(f, "%d", &P_imag->N_states) != 1) {
                    perror("The number of states of OUTPUT_IMAG is undetermined");
                    local_error = 1;
                }
                if (P->N_states > P_imag->N_states)
                    printf("WARNING: The number of states for the superposition of the initial wave function is "
                           "larger than the number of states computed in OUTPUT_IMAG");
            } else if (strcmp(key, "E") == 0 || strcmp(key, "DE") == 0) {
                for (j = 0; j < P_imag->N_states; j++)
                    if (fscanf(f, "%31s", key) != 1) local_error = 1;
            } else if (strcmp(key, "NR") == 0) {
                for (j = 0; j < P->dim; j++) {
                    long long v;
                    if (fscanf(f, "%lld", &v) != 1) {
                        printf("The number of grid points of OUTPUT_IMAG is not completely determined");
                        local_error = 1;
                    } else P_imag->Nr_global[j] = v;
                }
            } else if (strcmp(key, "DR") == 0) {
                for (j = 0; j < P->dim; j++)
                    if (fscanf(f, "%lf", &P->dr[j]) != 1) {
                        perror("The grid space of OUPUT_IMAG is not completely determined");
                        local_error = 1;
                    }
            } else if (strcmp(key, "POTENTIAL") == 0) {
                if (fscanf(f, "%19s", P->potential) != 1) {
                    perror("The potential name of OUPUT_IMAG is undetermined");
                    local_error = 1;
                }
            } else if (strcmp(key, "SOFTCORE") == 0) {  /* header line of older files of this code */
                if (fscanf(f, "%lf", &P->softcore) != 1) local_error = 1;
            } else {
                printf("\nUnknown parameter: %s\n", key);
                perror("Error with params.real");
                local_error = 1;
            }
        }
        fclose(f);
    } else {
        fprintf(stderr, "Rank 0 Error: Failed to re-open main parameter file for full read.\n");
        local_error = 1;
    }
    for (j = 0; j < dim; j++) P_imag->dr[j] = P->dr[j];
    {   /* soft-core extension: the side file written next to OUTPUT_IMAG */
        char side[MAX_OUTPUT_FILES_CHAR + 16];
        softcore_side_file(P->output_imag, side, sizeof side);
        FILE *s = fopen(side, "r");
        if (s) {
            if (fscanf(s, "%31s", key) != 1 || strcmp(key, "SOFTCORE") != 0 || fscanf(s, "%lf", &P->softcore) != 1) {
                printf("\nMalformed soft-core side file: %s\n", side);
                local_error = 1;
            }
            fclose(s);
        }
    }

    /* full read of params.real (:468-706) */
    f = fopen(filename, "r");
    if (f) {
        while (fscanf(f, "%31s", key) != EOF) {
            if (strcmp(key, "OUTPUT_IMAG") == 0) {
                Expected_arguments(f, "OUTPUT_IMAG", 1, Known);
                if (fscanf(f, "%29s", key) != 1) {
                    perror("The name of the OUPUT_IMAG is undetermined");
                    qnd_abort();
                }
            } else if (strcmp(key, "N_STATES") == 0) {
                Expected_arguments(f, "N_STATES", 1, Known);
                if (fscanf(f, "%d", &j) != 1) {
                    perror("The number of states is undetermined");
                    qnd_abort();
                }
            } else if (strcmp(key, "NR") == 0) {
                Expected_arguments(f, "NR", P->dim, Known);
                for (j = 0; j < P->dim; j++) {
                    long long v;
                    if (fscanf(f, "%lld", &v) != 1) {
                        perror("The number of grid points is not completely determined");
                        qnd_abort();
                    }
                    P->Nr_global[j] = v;
                    if (!is_pow2(P->Nr_global[j])) {
                        printf("\nThe number of points on the grid must be a power of 2\n");
                        perror("Error with grid points number");
                        qnd_abort();
                    }
                    if (P->Nr_global[j] < P_imag->Nr_global[j])
                        printf("WARNING: The number of points on the grid is smaller than for the imaginary time "
                               "propagation\n");
                }
            } else if (strcmp(key, "J_STATES") == 0) {
                Expected_arguments(f, "J_STATES", P->N_states, Known);
                for (j = 0; j < P->N_states; j++) {
                    if (fscanf(f, "%d", &P->j_states[j]) != 1) {
                        perror("An ith state is undetermined");
                        local_error = 1;
                    }
                    if (P->j_states[j] > P_imag->N_states) {
                        printf("\nAn initial eigenstate is not computed in the imaginary time propagation\n");
                        perror("Error with initial eigenstate");
                        local_error = 1;
                    }
                }
            } else if (strcmp(key, "COEFFS_STATES") == 0) {
                Expected_arguments(f, "COEFFS_STATES", P->N_states, Known);
                for (j = 0; j < P->N_states; j++) {
                    if (fscanf(f, "%31s", key) != 1) {
                        perror("The coefficient of a, ith state is undetermined");
                        local_error = 1;
                    }
                    parse_complex(key, &P->coeffs_states[2 * j], &P->coeffs_states[2 * j + 1]);
                }
            } else if (strcmp(key, "CAP") == 0) {
                Expected_arguments(f, "CAP", 2, Known);
                if (fscanf(f, "%lf%lf", &P->cap[0], &P->cap[1]) != 2) {
                    perror("The CAP parameters are undetermined (there must be 2 values)");
                    local_error = 1;
                }
                P->cap[0] = fabs(P->cap[0]);
                P->cap[1] = fabs(P->cap[1]);
            } else if (strcmp(key, "N_OBSERVABLES") == 0) {
                Expected_arguments(f, "N_OBSERVABLES", 1, Known);
                if (fscanf(f, "%d", &j) != 1) {
                    perror("The number of observables is undetermined");
                    local_error = 1;
                }
            } else if (strcmp(key, "DT_SAVE") == 0) {
                Expected_arguments(f, "DT_SAVE", 1, Known);
                read_doubles(f, &P->dt_save, 1, "The time-step for the observables is undetermined", &local_error);
            } else if (strcmp(key, "P_MAX") == 0) {
                Expected_arguments(f, "P_MAX", P->dim, Known);
                read_doubles(f, P->pr_max, P->dim, "The maximum number of momentum coordinate is undetermined",
                             &local_error);
            } else if (strcmp(key, "R_MASK") == 0) {
                Expected_arguments(f, "R_MASK", 2, Known);
                if (fscanf(f, "%lf%lf", &P->R_mask[0], &P->R_mask[1]) != 2) {
                    perror("The parameters of the mask are undetermined");
                    local_error = 1;
                }
            } else if (strcmp(key, "DE") == 0) {
                Expected_arguments(f, "DE", 1, Known);
                read_doubles(f, &P->de, 1, "The energy step for the PED is undetermined", &local_error);
            } else if (strcmp(key, "NE") == 0) {
                Expected_arguments(f, "NE", 1, Known);
                if (fscanf(f, "%d", &P->Ne) != 1) {
                    perror("The number of energy step for the PED is undetermined");
                    local_error = 1;
                }
            } else if (strcmp(key, "OBSERVABLES") == 0) {
                Expected_arguments(f, "OBSERVABLES", P->N_observables, Known);
                for (j = 0; j < P->N_observables; j++)
                    if (fscanf(f, "%31s", P->observables[j]) != 1) {
                        perror("Not enough observable names found.");
                        local_error = 1;
                    }
            } else if (strcmp(key, "DT") == 0) {
                Expected_arguments(f, "DT", 1, Known);
                read_doubles(f, &P->dt, 1, "The time step is undetermined", &local_error);
            } else if (strcmp(key, "TINTEGRATION") == 0) {
                Expected_arguments(f, "TINTEGRATION", 1, Known);
                read_doubles(f, &P->Tintegration, 1, "The integration time is undetermined", &local_error);
            } else if (strcmp(key, "INTEGRATOR") == 0) {
                Expected_arguments(f, "INTEGRATOR", 1, Known);
                if (fscanf(f, "%9s", P->integrator) != 1) {
                    perror("The integrator is undetermined");
                    local_error = 1;
                }
            } else if (strcmp(key, "INTENSITY") == 0) {
                Expected_arguments(f, "INTENSITY", F->N_modes, Known);
                read_doubles(f, F->intensity, F->N_modes, "One field intensity is undetermined", &local_error);
            } else if (strcmp(key, "N_MODES") == 0) {
                Expected_arguments(f, "N_MODES", 1, Known);
                if (fscanf(f, "%d", &j) != 1) {
                    perror("The number of modes is undetermined");
                    local_error = 1;
                }
            } else if (strcmp(key, "LAMBDA") == 0) {
                Expected_arguments(f, "LAMBDA", F->N_modes, Known);
                read_doubles(f, F->lambda, F->N_modes, "One field wavelength is undetermined", &local_error);
            } else if (strcmp(key, "ELLIPTICITY") == 0) {
                Expected_arguments(f, "ELLIPTICITY", F->N_modes, Known);
                read_doubles(f, F->ellipticity, F->N_modes, "One field ellipticity is undetermined", &local_error);
            } else if (strcmp(key, "CEP") == 0) {
                Expected_arguments(f, "CEP", F->N_modes, Known);
                read_doubles(f, F->cep, F->N_modes, "One field CEP is undetermined", &local_error);
            } else if (strcmp(key, "DELAY") == 0) {
                Expected_arguments(f, "DELAY", F->N_modes, Known);
                for (j = 0; j < F->N_modes; j++) {
                    if (fscanf(f, "%lf", &F->delay[j]) != 1) {
                        perror("One field delay is undetermined");
                        local_error = 1;
                    }
                    if (F->delay[j] < 0) {
                        perror("Delays must be positive (the time integration starts at 0 always)");
                        local_error = 1;
                    }
                }
            } else if (strcmp(key, "DURATION") == 0) {
                Expected_arguments(f, "DURATION", F->N_modes, Known);
                read_doubles(f, F->duration, F->N_modes, "One field duration is undetermined", &local_error);
            } else if (strcmp(key, "GAUGE") == 0) {
                Expected_arguments(f, "GAUGE", 1, Known);
                if (fscanf(f, "%9s", F->gauge) != 1) {
                    perror("One field gauge is undetermined");
                    local_error = 1;
                }
            } else if (strcmp(key, "ENVELOPE") == 0) {
                Expected_arguments(f, "ENVELOPE", F->N_modes, Known);
                for (j = 0; j < F->N_modes; j++)
                    if (fscanf(f, "%9s", F->envelope[j]) != 1) {
                        perror("One field envelope is undetermined");
                        local_error = 1;
                    }
            } else if (strcmp(key, "MASK") == 0) {
                Expected_arguments(f, "MASK", 1, Known);
                read_doubles(f, &P->mask_frac, 1, "The mask fraction is undetermined", &local_error);
            } else if (strcmp(key, "PRECISION") == 0) {
                Expected_arguments(f, "PRECISION", 1, Known);
                if (fscanf(f, "%63s", val) != 1) local_error = 1;
                else P->precision = parse_precision(val);
            } else if (strcmp(key, "STEPS_PER_CALL") == 0) {
                Expected_arguments(f, "STEPS_PER_CALL", 1, Known);
                if (fscanf(f, "%d", &P->chunk) != 1 || P->chunk < 1) local_error = 1;
            } else {
                printf("\nUnknown parameter: %s\n", key);
                perror("Error with params.real");
                local_error = 1;
            }
        }
        fclose(f);
    } else {
        fprintf(stderr, "Rank 0 Error: Failed to open imaginary output file for parameter read.\n");
        local_error = 1;
    }
    if (local_error) {
        fprintf(stderr, "\n FATAL ERROR: Data arrays or complex fields failed to read. Aborting.\n");
        qnd_abort();
    }

    /* overrides, mode 0 only (:715-771) */
    for (j = 2; j < argc; j += 2) {
        if (j + 1 < argc) {
            const char *k = argv[j], *v = argv[j + 1];
            if (strcmp(k, "DT") == 0) P->dt = atof(v);
            else if (strcmp(k, "TINTEGRATION") == 0) P->Tintegration = atof(v);
            else if (strcmp(k, "NE") == 0) P->Ne = atoi(v);
            else if (strcmp(k, "DE") == 0) P->de = atof(v);
            else if (strcmp(k, "INTENSITY") == 0) F->intensity[0] = atof(v);
            else if (strcmp(k, "ELLIPTICITY") == 0) F->ellipticity[0] = atof(v);
            else if (strcmp(k, "LAMBDA") == 0) F->lambda[0] = atof(v);
            else if (strcmp(k, "CEP") == 0) F->cep[0] = atof(v);
            else if (strcmp(k, "DELAY") == 0) F->delay[0] = atof(v);
            else if (strcmp(k, "DURATION") == 0) F->duration[0] = atof(v);
            else {
                fprintf(stderr, "Rank 0 ERROR: Unrecognized parameter key '%s'. Check spelling.\n", k);
                qnd_abort();
            }
        } else {
            fprintf(stderr, "Rank 0 ERROR: Parameter key '%s' found without a corresponding value.\n", argv[j]);
            qnd_abort();
        }
    }

    /* normalise the superposition coefficients (:801-804) */
    double norm2 = 0.;
    for (j = 0; j < P->N_states; j++) {
        double m = hypot(P->coeffs_states[2 * j], P->coeffs_states[2 * j + 1]);
        norm2 += m * m;
    }
    for (j = 0; j < P->N_states; j++) {
        P->coeffs_states[2 * j] /= sqrt(norm2);
        P->coeffs_states[2 * j + 1] /= sqrt(norm2);
    }
    /* laser constants (:807-810) */
    for (j = 0; j < F->N_modes; j++) {
        F->amplitude[j] = 5.33803e-9 * sqrt(F->intensity[j]);
        F->frequency[j] = 45.5637639947958 / (double)F->lambda[j];
    }

    printf("> For real-time propagation:\n");
    printf("  DIMENSION\t\t%d\n", P->dim);
    printf("  NR\t\t");
    for (j = 0; j < P->dim; j++) printf("\t%lld", (long long)P->Nr_global[j]);
    printf("\n  DR\t\t");
    for (j = 0; j < P->dim; j++) printf("\t%16.15f", P->dr[j]);
    printf("\n  N_STATES\t\t%d\n  J_STATES\t", P->N_states);
    for (j = 0; j < P->N_states; j++) printf("\t%d", P->j_states[j]);
    printf("\n  COEFFS_STATES\t");
    for (j = 0; j < P->N_states; j++)
        printf("\t%16.15f+I*%16.15f", P->coeffs_states[2 * j], P->coeffs_states[2 * j + 1]);
    printf("\n  POTENTIAL\t\t%s\n", P->potential);
    printf("  CAP\t\t\t%16.15f\t%16.15f\n", P->cap[0], P->cap[1]);
    printf("  INTEGRATOR\t\t%s\n  DT\t\t\t%16.15f\n  TINTEGRATION\t\t%16.15f\n\n", P->integrator, P->dt,
           P->Tintegration);
    printf("> Field parameters:\n  GAUGE\t\t\t%s\n  N_MODES\t\t%d\n  ENVELOPE\t", F->gauge, F->N_modes);
    for (j = 0; j < F->N_modes; j++) printf("\t%s", F->envelope[j]);
    printf("\n  DURATION\t");
    for (j = 0; j < F->N_modes; j++) printf("\t%16.15f", F->duration[j]);
    printf("\n  INTENSITY\t");
    for (j = 0; j < F->N_modes; j++) printf("\t%6.5e", F->intensity[j]);
    printf("\n  LAMBDA\t");
    for (j = 0; j < F->N_modes; j++) printf("\t%16.15f", F->lambda[j]);
    printf("\n  ELLIPTICITY\t");
    for (j = 0; j < F->N_modes; j++) printf("\t%16.15f", F->ellipticity[j]);
    printf("\n  CEP\t\t");
    for (j = 0; j < F->N_modes; j++) printf("\t%16.15f", F->cep[j]);
    printf("\n  DELAY\t\t");
    for (j = 0; j < F->N_modes; j++) printf("\t%16.15f", F->delay[j]);
    printf("\n\n> Observables:\n  N_OBSERVABLES\t\t%d\n  OBSERVABLES\t", P->N_observables);
    for (j = 0; j < P->N_observables; j++) printf("\t%s", P->observables[j]);
    printf("\n  DT_SAVE\t\t%16.15f\n  P_MAX\t\t", P->dt_save);
    for (j = 0; j < P->dim; j++) printf("\t%16.15f", P->pr_max[j]);
    printf("\n  R_MASK\t\t%16.15f\t%16.15f\n  DE\t\t\t%16.15f\n  NE\t\t\t%d\n\n", P->R_mask[0], P->R_mask[1],
           P->de, P->Ne);
    printf("> For the imaginary-time propagation's data taken from '%s':\n", P->output_imag);
    printf("  OUTPUT_IMAG\t\t%s\n  DIMENSION\t\t%d\n  NR\t\t", P->output_imag, P_imag->dim);
    for (j = 0; j < P->dim; j++) printf("\t%lld", (long long)P_imag->Nr_global[j]);
    printf("\n  N_STATES\t\t%d\n", P_imag->N_states);
    fflush(stdout);

    /* time vector (:886-895) */
    P->N_time = (int64_t)(P->Tintegration / (double)P->dt);
    if (P->N_time < 0) P->N_time = 0;
    P->time = (double *)malloc((P->N_time > 0 ? P->N_time : 1) * sizeof(double));
    for (int64_t i = 0; i < P->N_time; i++) P->time[i] = i * P->dt;
}

void Finalize_parameters(Parameters *P)
{
    if (!P) return;
    free(P->Nr_global);
    free(P->dr);
    free(P->pr_max);
    free(P->time);
    free(P->j_states);
    free(P->coeffs_states);
    if (P->observables)
        for (int j = 0; j < P->N_observables; j++) free(P->observables[j]);
    free(P->observables);
    memset(P, 0, sizeof *P);
}

void Finalize_field_vectors(Field *F)
{
    if (!F) return;
    free(F->intensity);
    free(F->lambda);
    free(F->cep);
    free(F->ellipticity);
    free(F->duration);
    free(F->delay);
    free(F->frequency);
    free(F->amplitude);
    if (F->envelope)
        for (int j = 0; j < F->N_modes; j++) free(F->envelope[j]);
    free(F->envelope);
    free(F->E);
    free(F->A);
    memset(F, 0, sizeof *F);
}

/* ------------------------------------------------------------------ */
/* laser field (src/parameters.c:899-949)                               */
/* ------------------------------------------------------------------ */
void Field_vectors(const Parameters *P, const Field *F, double t, double *A, double *E)
{
    double f = 0., df = 0.;
    for (int j = 0; j < P->dim; j++) {
        E[j] = 0.;
        A[j] = 0.;
    }
    for (int j = 0; j < F->N_modes; j++) {
        const double w = F->frequency[j], F0 = F->amplitude[j], xi = F->ellipticity[j];
        const double tau = F->duration[j], ph = F->cep[j], ts = t - F->delay[j];
        const char *env = F->envelope[j];
        if (strcmp(env, "constant") == 0) {
            f = 1.;
            df = 0.;
        } else if (strcmp(env, "trapz") == 0) {
            const double tr = 2. * 2. * M_PI / w;   /* two-cycle ramps */
            const double up = (ts >= 0. && ts <= tr), flat = (ts > tr && ts < (tau - tr));
            const double down = (ts >= tau - tr && ts <= tau);
            f = pow(sin(.5 * M_PI * ts / tr), 4) * up + flat + pow(sin(.5 * M_PI * (ts - tau) / tr), 4) * down;
            df = M_PI / tr *
                 (pow(sin(.5 * M_PI * ts / tr), 2) * sin(M_PI * ts / tr) * up +
                  pow(sin(.5 * M_PI * (ts - tau) / tr), 2) * sin(M_PI * (ts - tau) / tr) * down);
        } else if (strcmp(env, "cos2") == 0) {
            const double in = (ts >= 0. && ts <= tau);
            f = pow(sin(M_PI * ts / tau), 2) * in;
            df = M_PI * sin(2. * M_PI * ts / tau) / tau * in;
        } else if (strcmp(env, "cos4") == 0) {
            const double in = (ts >= 0. && ts <= tau);
            f = pow(sin(M_PI * ts / tau), 4) * in;
            df = 4. * M_PI * cos(M_PI * ts / tau) * pow(sin(M_PI * ts / tau), 3) / tau * in;
        } else {
            printf("\nUnknown envelope: %s\n", env);
            perror("Error with the envelope name");
            qnd_abort();
        }
        /* A = F0 f (cos, xi sin)(w t + phi) / w,  E = -dA/dt */
        E[0] += -F0 * (df * cos(w * ts + ph) - f * w * sin(w * ts + ph)) / w;
        if (P->dim >= 2) E[1] += -F0 * xi * (df * sin(w * ts + ph) + f * w * cos(w * ts + ph)) / w;
        A[0] += F0 * f * cos(w * ts + ph) / w;
        if (P->dim >= 2) A[1] += F0 * xi * f * sin(w * ts + ph) / w;
    }
}

/* src/integrators.c:13-71 */
int integrator_coefficients(const char *name, double dt, double *as, double *as_scaled)
{
    static const double bm4[6] = {.0792036964311957, .1303114101821663, .2228614958676077,
                                  -.3667132690474257, .3246481886897062, .1096884778767498};
    static const double bm6[10] = {.050262764400392, .098553683500650, .314960616927694, -.447346482695478,
                                   .492426372489876, -.425118767797691, .237063913978122, .195602488600053,
                                   .346358189850727, -.362762779254345};
    static const double rkn4[6] = {.082984406417405, .162314550766866, .233995250731502,
                                   .370877414979578, -.409933719901926, .059762097006575};
    int s;
    for (int j = 0; j < 20; j++) as[j] = as_scaled[j] = 0.;
    if (strcmp(name, "Verlet") == 0) {
        s = 1;
        as[0] = .5;
    } else if (strcmp(name, "BM4") == 0) {
        s = 6;
        memcpy(as, bm4, sizeof bm4);
    } else if (strcmp(name, "BM6") == 0) {
        s = 10;
        memcpy(as, bm6, sizeof bm6);
    } else if (strcmp(name, "RKN4") == 0) {
        s = 6;
        memcpy(as, rkn4, sizeof rkn4);
    } else {
        return -1;
    }
    for (int j = 0; j < s; j++) as[s + j] = as[s - 1 - j];
    for (int j = 0; j < 2 * s; j++) as_scaled[j] = as[j] * dt;
    return s;
}

/* src/integrators.c:74-129: stage-time field tables and the two ASCII files */
void Initialize_field_vectors(const Parameters *P, Field *F, const double *as_scaled, int s)
{
    const int dim = P->dim;
    const size_t n = (size_t)dim * (size_t)P->N_time * (size_t)s;
    F->E = (double *)calloc(n > 0 ? n : 1, sizeof(double));
    F->A = (double *)calloc(n > 0 ? n : 1, sizeof(double));
    if (!F->E || !F->A) {
        fprintf(stderr, "ERROR: memory allocation failed for F->E or F->A in Initialize_parameters.\n");
        qnd_abort();
    }
    for (int64_t i = 0; i < P->N_time; i++) {
        double t = P->time[i] + as_scaled[0];
        const size_t ind = (size_t)i * dim * s;
        for (int j = 0; j < s; j++) {
            Field_vectors(P, F, t, &F->A[ind + j * dim], &F->E[ind + j * dim]);
            if (j < s - 1) t += as_scaled[2 * j + 1] + as_scaled[2 * j + 2];
        }
    }
    FILE *fe = fopen("electric_field.real", "w"), *fa = fopen("vector_potential.real", "w");
    if (!fe || !fa) {
        perror("Error opening the files of the vector potential and the laser electric field.");
        qnd_abort();
    }
    for (int64_t i = 0; i < P->N_time; i++) {
        double t = P->time[i] + as_scaled[0];
        const size_t ind = (size_t)i * dim * s;
        for (int j = 0; j < s; j++) {
            fprintf(fe, "%16.15f ", t);
            fprintf(fa, "%16.15f ", t);
            for (int k = 0; k < dim; k++) {
                fprintf(fe, "%16.15f ", F->E[ind + j * dim + k]);
                fprintf(fa, "%16.15f ", F->A[ind + j * dim + k]);
            }
            fprintf(fe, "\n");
            fprintf(fa, "\n");
            if (j < s - 1) t += as_scaled[2 * j + 1] + as_scaled[2 * j + 2];
        }
    }
    fclose(fe);
    fclose(fa);
}

/* SplitMix64 -> uniform (0,1] -> Box-Muller normals (extension) */
static uint64_t splitmix64(uint64_t *st)
{
    uint64_t z = (*st += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void gaussian_stream(uint64_t seed, int64_t n, double *out)
{
    uint64_t st = seed;
    for (int64_t i = 0; i < n; i += 2) {
        const double u1 = ((double)(splitmix64(&st) >> 11) + 1.0) * 0x1.0p-53;
        const double u2 = ((double)(splitmix64(&st) >> 11) + 1.0) * 0x1.0p-53;
        const double rad = sqrt(-2.0 * log(u1));
        out[i] = rad * cos(2.0 * M_PI * u2);
        if (i + 1 < n) out[i + 1] = rad * sin(2.0 * M_PI * u2);
    }
}

int observable_code(const char *name)
{
    if (strcmp(name, "Lz") == 0) return 1;
    if (strcmp(name, "energy") == 0) return 2;
    if (strcmp(name, "dipole_r") == 0) return 3;
    if (strcmp(name, "dipole_v") == 0) return 4;
    if (strcmp(name, "norm") == 0) return 5;
    if (strcmp(name, "dipole_a") == 0) return 6;
    return 0;
}

int observable_width(const char *name, int dim)
{
    switch (observable_code(name)) {
    case 1: case 2: case 5: return 1;
    case 3: case 4: case 6: return dim;
    default: return 0;
    }
}
